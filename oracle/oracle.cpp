// CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h for scope).
//
// Scalar, one path at a time, depth-major as the reference's bounce loop.
// Strict IEEE: built with -O2 -ffp-contract=off, no fast-math.
#include "oracle.h"

#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../raymond_amd/csrc/device/bvh2l.h"
#include "../raymond_amd/csrc/device/isect.h"
#include "../raymond_amd/csrc/device/material_prelude.h"
#include "../raymond_amd/csrc/device/camera.h"

namespace {

typedef void (*MatFn)(uint32_t, const rm::SceneView*, rm::ShadingContext*);

struct MatTable {
    MatFn surface = nullptr;
    MatFn light = nullptr;
};
thread_local const MatTable* t_mat = nullptr;

}  // namespace

// The material hooks the shared device library calls (device/shading.hpp:11-12)
namespace rm {
inline void shade_surface(uint32_t i, const SceneView& c, ShadingContext& s) { t_mat->surface(i, &c, &s); }
inline void shade_light(uint32_t i, const SceneView& c, ShadingContext& s) { t_mat->light(i, &c, &s); }
}  // namespace rm

using rm::vec2;
using rm::vec3;

namespace {

struct WTri {
    vec3 v0, v1, v2;
    uint32_t inst, prim;
};

struct BNode {
    float lo[3], hi[3];
    int32_t left;    // -1: leaf
    int32_t right;
    uint32_t first, count;
};

struct Hit {
    float t = -1.0f;
    uint32_t inst = 0xffffffffu, prim = 0xffffffffu;
    float b0 = 0.0f, b1 = 0.0f;
};

}  // namespace

struct or_scene {
    std::vector<float> vertices, normals, texcoords;
    std::vector<uint32_t> indices;
    std::vector<uint16_t> materials;
    std::vector<rm::PerInstance> instances;
    std::vector<rm::AreaLight> area;
    std::vector<rm::PointLight> point;
    std::vector<rm::SunLight> sun;
    std::vector<rm::SpotLight> spot;
    std::vector<rm::ShapeLight> shape;
    std::vector<float> light_faces;
    std::vector<float> env_pdfs, env_mip;
    int env_res = 2048;
    std::vector<std::vector<float>> tex_pixels;  // texture slots of the material library
    std::vector<rm::Texture> tex;
    rm::SceneView view{};
    std::vector<WTri> tris;
    std::vector<BNode> bvh;
    std::vector<uint32_t> bvh_idx;
    std::vector<uint32_t> inst_faces;   // face count of each instance
    // two-level scenes: world-to-object rows per instance (or_scene_set_two_level)
    bool two_level = false;
    std::vector<rm::TlasInstance> tlas;
    void* so = nullptr;
    MatTable mat;
};

namespace {

// ---------------------------------------------------------------------------
// Scene restatement

rm::mat4 mat4_of(const float* m) {
    rm::mat4 r;
    r.c0 = rm::vec4(m[0], m[1], m[2], m[3]);
    r.c1 = rm::vec4(m[4], m[5], m[6], m[7]);
    r.c2 = rm::vec4(m[8], m[9], m[10], m[11]);
    r.c3 = rm::vec4(m[12], m[13], m[14], m[15]);
    return r;
}

// inner3x3.inverse.transpose (EntityBuilder.swift:98), in double via the adjugate
rm::mat3 inverse_transpose(const rm::mat4& M) {
    double a[3][3];  // a[row][col]
    const rm::vec4 c[3] = {M.c0, M.c1, M.c2};
    for (int j = 0; j < 3; j++) {
        a[0][j] = c[j].x;
        a[1][j] = c[j].y;
        a[2][j] = c[j].z;
    }
    double cof[3][3];
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++)
            cof[r][k] = a[(r + 1) % 3][(k + 1) % 3] * a[(r + 2) % 3][(k + 2) % 3] -
                        a[(r + 1) % 3][(k + 2) % 3] * a[(r + 2) % 3][(k + 1) % 3];
    const double det = a[0][0] * cof[0][0] + a[0][1] * cof[0][1] + a[0][2] * cof[0][2];
    double inv[3][3];  // inverse = adjugate / det, adjugate = cofactor^T
    for (int r = 0; r < 3; r++)
        for (int k = 0; k < 3; k++) inv[r][k] = cof[k][r] / det;
    rm::mat3 out;  // transpose: out[row r][col k] = inv[k][r]
    rm::vec3* cols[3] = {&out.c0, &out.c1, &out.c2};
    for (int k = 0; k < 3; k++) *cols[k] = vec3((float)inv[k][0], (float)inv[k][1], (float)inv[k][2]);
    return out;
}

// host/lights/distribution.m:7-39
float light_distribution(const rm::mat3& M, const uint32_t* idx, const float* verts, uint32_t n, float* out) {
    double acc = 0.0;
    for (uint32_t i = 0; i < n; i++) {
        const float* p0 = verts + 3 * idx[3 * i];
        const float* p1 = verts + 3 * idx[3 * i + 1];
        const float* p2 = verts + 3 * idx[3 * i + 2];
        const vec3 v0(p0[0], p0[1], p0[2]), v1(p1[0], p1[1], p1[2]), v2(p2[0], p2[1], p2[2]);
        const vec3 tn = rm::cross(M * (v2 - v0), M * (v2 - v1));
        const float area = rm::length(tn) / 2.0f;
        out[i] = (float)(acc += area);
    }
    for (uint32_t i = 0; i < n; i++) out[i] = (float)((double)out[i] / acc);
    return (float)acc;
}

// ---------------------------------------------------------------------------
// Oracle BVH2: median split on the widest centroid axis, leaves <= 4.

void bvh_box(const or_scene& s, uint32_t b, uint32_t e, BNode& n) {
    for (int a = 0; a < 3; a++) {
        n.lo[a] = INFINITY;
        n.hi[a] = -INFINITY;
    }
    for (uint32_t i = b; i < e; i++) {
        const WTri& t = s.tris[s.bvh_idx[i]];
        const vec3 vs[3] = {t.v0, t.v1, t.v2};
        for (const vec3& v : vs)
            for (int a = 0; a < 3; a++) {
                n.lo[a] = std::min(n.lo[a], rm::comp(v, a));
                n.hi[a] = std::max(n.hi[a], rm::comp(v, a));
            }
    }
    // one ulp outward: a ray parallel to an axis that lies in a face of the
    // geometry's bounds then starts strictly inside the slab, where the slab
    // test's 0 * (1 / 1e-30) would otherwise close it at t = 0
    for (int a = 0; a < 3; a++) {
        n.lo[a] = std::nextafter(n.lo[a], -INFINITY);
        n.hi[a] = std::nextafter(n.hi[a], INFINITY);
    }
}

int bvh_build(or_scene& s, uint32_t b, uint32_t e) {
    const int id = (int)s.bvh.size();
    s.bvh.emplace_back();
    BNode n{};
    bvh_box(s, b, e, n);
    if (e - b <= 4) {
        n.left = n.right = -1;
        n.first = b;
        n.count = e - b;
        s.bvh[id] = n;
        return id;
    }
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    auto cent = [&](uint32_t t, int a) {
        const WTri& w = s.tris[t];
        return (rm::comp(w.v0, a) + rm::comp(w.v1, a) + rm::comp(w.v2, a)) / 3.0f;
    };
    for (uint32_t i = b; i < e; i++)
        for (int a = 0; a < 3; a++) {
            clo[a] = std::min(clo[a], cent(s.bvh_idx[i], a));
            chi[a] = std::max(chi[a], cent(s.bvh_idx[i], a));
        }
    int axis = 0;
    for (int a = 1; a < 3; a++)
        if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
    const uint32_t mid = b + (e - b) / 2;
    std::nth_element(s.bvh_idx.begin() + b, s.bvh_idx.begin() + mid, s.bvh_idx.begin() + e,
                     [&](uint32_t x, uint32_t y) {
                         const float cx = cent(x, axis), cy = cent(y, axis);
                         return cx < cy || (cx == cy && x < y);
                     });
    const int l = bvh_build(s, b, mid);
    const int r = bvh_build(s, mid, e);
    n.left = l;
    n.right = r;
    s.bvh[id] = n;
    return id;
}

bool test_tri(const WTri& t, vec3 o, const rm::RayShear& sh, float tmin, float tmax, Hit& best) {
    float tt, b0, b1;
    if (!rm::intersect_triangle(o, sh, t.v0, t.v1, t.v2, tmin, tmax, tt, b0, b1)) return false;
    if (best.t >= 0.0f && !rm::hit_before(tt, t.inst, t.prim, best.t, best.inst, best.prim)) return false;
    best.t = tt;
    best.inst = t.inst;
    best.prim = t.prim;
    best.b0 = b0;
    best.b1 = b1;
    return true;
}

// Two-level scenes (the instance acceleration structure of the reference,
// EntityBuilder.swift:146-165): every instance, the world ray mapped into its
// object space by its world-to-object rows, every triangle of its shape in
// object space.  Exhaustive, so it checks the product's culling too.
Hit trace_two_level(const or_scene& s, vec3 wo, vec3 wd, float tmin, float tmax, bool any) {
    Hit best;
    for (uint32_t k = 0; k < (uint32_t)s.instances.size(); k++) {
        const rm::TlasInstance& ti = s.tlas[k];
        const vec3 o = rm::tlas_point(ti, wo), d = rm::tlas_vector(ti, wd);
        const rm::RayShear sh = rm::make_shear(d);
        const rm::PerInstance& pi = s.instances[k];
        for (uint32_t f = 0; f < s.inst_faces[k]; f++) {
            const uint32_t* ix = &s.indices[3 * ((size_t)pi.face_offset + f)];
            WTri t;
            const float* p0 = &s.vertices[3 * ((size_t)pi.vertex_offset + ix[0])];
            const float* p1 = &s.vertices[3 * ((size_t)pi.vertex_offset + ix[1])];
            const float* p2 = &s.vertices[3 * ((size_t)pi.vertex_offset + ix[2])];
            t.v0 = vec3(p0[0], p0[1], p0[2]);
            t.v1 = vec3(p1[0], p1[1], p1[2]);
            t.v2 = vec3(p2[0], p2[1], p2[2]);
            t.inst = k;
            t.prim = f;
            if (test_tri(t, o, sh, tmin, tmax, best) && any) return best;
        }
    }
    return best;
}

// closest hit (or any hit) with t in (tmin, tmax)
Hit trace(const or_scene& s, vec3 o, vec3 d, float tmin, float tmax, bool any, bool brute) {
    if (s.two_level) return trace_two_level(s, o, d, tmin, tmax, any);
    Hit best;
    const rm::RayShear sh = rm::make_shear(d);
    if (brute) {
        for (const WTri& t : s.tris)
            if (test_tri(t, o, sh, tmin, tmax, best) && any) return best;
        return best;
    }
    if (s.bvh.empty()) return best;
    const vec3 inv = rm::safe_inv_dir(d);
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const BNode& n = s.bvh[stack[--sp]];
        float tn;
        const float lim = best.t >= 0.0f ? best.t : tmax;
        if (!rm::box_hit(vec3(n.lo[0], n.lo[1], n.lo[2]), vec3(n.hi[0], n.hi[1], n.hi[2]), o, inv, tmin, lim, tn))
            continue;
        if (n.left < 0) {
            for (uint32_t i = 0; i < n.count; i++)
                if (test_tri(s.tris[s.bvh_idx[n.first + i]], o, sh, tmin, tmax, best) && any) return best;
            continue;
        }
        const BNode& L = s.bvh[n.left];
        const BNode& R = s.bvh[n.right];
        float tl, tr;
        const bool hl = rm::box_hit(vec3(L.lo[0], L.lo[1], L.lo[2]), vec3(L.hi[0], L.hi[1], L.hi[2]), o, inv, tmin, lim, tl);
        const bool hr = rm::box_hit(vec3(R.lo[0], R.lo[1], R.lo[2]), vec3(R.hi[0], R.hi[1], R.hi[2]), o, inv, tmin, lim, tr);
        if (hl && hr) {
            if (tl <= tr) {
                stack[sp++] = n.right;
                stack[sp++] = n.left;
            } else {
                stack[sp++] = n.left;
                stack[sp++] = n.right;
            }
        } else if (hl) {
            stack[sp++] = n.left;
        } else if (hr) {
            stack[sp++] = n.right;
        }
    }
    return best;
}

template <typename F>
void parallel_for(uint32_t n, int threads, F f) {
    int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt == 1 || n < 64) {
        for (uint32_t i = 0; i < n; i++) f(i, 0);
        return;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
        pool.emplace_back([&, t] {
            for (;;) {
                const uint32_t b = next.fetch_add(64);
                if (b >= n) break;
                const uint32_t e = std::min(n, b + 64);
                for (uint32_t i = b; i < e; i++) f(i, t);
            }
        });
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Kernels, one path at a time

// power heuristic (intersection.metal:14-20)
float mis_weight(float pdf, float other) {
    pdf *= pdf;
    if (rm::isinf_(pdf)) return 1.0f;
    other *= other;
    return pdf / (pdf + other);
}

// device/utils/math.hpp:73-119
vec3 valid_reflection(vec3 Ng, vec3 I, vec3 N) {
    using namespace rm;
    const vec3 R = 2.0f * dot(N, I) * N - I;
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
        const float R1 = 2.0f * (N1.x * Ix + N1.y * Iz) * N1.y - Iz;
        const float R2 = 2.0f * (N2.x * Ix + N2.y * Iz) * N2.y - Iz;
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

struct Ray {  // bridge/Ray.hpp:17-29
    vec3 origin;
    float min_distance;
    vec3 direction;
    float max_distance;
    rm::Prng prng;
    vec3 weight;
    uint32_t x, y;
    uint32_t depth;
    uint32_t flags;
    float bsdf_pdf;
};

struct ShadowRay {  // bridge/Ray.hpp:31-39
    vec3 origin;
    float min_distance;
    vec3 direction;
    float max_distance;
    vec3 weight;
};

// intersection.metal:84-87, shadow.metal:24-27: RGB += c, alpha += 1
void add(float* p, vec3 c) {
    p[0] += c.x;
    p[1] += c.y;
    p[2] += c.z;
    p[3] += 1.0f;
}

// generate.metal:88-91 (8x8 threadgroups; smaller at the right/bottom edges)
uint32_t block_linear_index(uint32_t x, uint32_t y, uint32_t W, uint32_t H) {
    const uint32_t tx = x % 8, ty = y % 8, gx = x / 8, gy = y / 8;
    const uint32_t awx = std::min(8u, W - 8 * gx), awy = std::min(8u, H - 8 * gy);
    return tx + ty * awx + gx * 8 * awy + gy * 8 * W;
}

struct PathCounts {
    uint64_t rays[33] = {0};
    uint64_t shadows[32] = {0};
};

// generateRays + depth loop of handleIntersections / handleShadowRays for one path
// px: the RGBA sums of pixel (x, y)
void render_path(const or_scene& S, const or_render_args& A, const rm::Lens& lens, uint32_t x, uint32_t y,
                 uint32_t frame, float* px, PathCounts& cnt) {
    using namespace rm;
    const SceneView& ctx = S.view;
    const uint32_t W = A.width, H = A.height;
    // generateRays (generate.metal:93-121)
    Ray ray;
    const uint32_t ray_index = block_linear_index(x, y, W, H);
    ray.prng = Prng(frame + 1, ray_index);
    ray.min_distance = ctx.camera.near_clip;
    ray.max_distance = ctx.camera.far_clip;
    ray.flags = RayCamera;
    ray.depth = 0;
    ray.bsdf_pdf = kInf;
    ray.x = x;
    ray.y = y;
    if (lens.num_surfaces > 0) {
        // a ray the lens stops is never committed (generate.metal:143-146)
        if (!rm::lens_camera_ray(ctx.camera, lens, x, y, W, H, ray.prng, ray.origin, ray.direction, ray.weight))
            return;
    } else {
        const vec2 jr = ray.prng.sample2d();
        const vec2 jittered((float)x + jr.x, (float)y + jr.y);
        const vec2 uv(+1.0f * ((jittered.x / (float)W + ctx.camera.shift.x) * 2.0f - 1.0f),
                      -1.0f * ((jittered.y / (float)H + ctx.camera.shift.y) * 2.0f - 1.0f));
        const float aspect = (float)H / (float)W;
        ray.origin = ctx.camera.transform.point(vec3(0.0f, 0.0f, 0.0f));
        ray.direction = normalize(ctx.camera.transform.vector(vec3(uv.x, uv.y * aspect, -ctx.camera.focal_length)));
        ray.weight = vec3(1.0f);
    }
    Hit isect = trace(S, ray.origin, ray.direction, ray.min_distance, ray.max_distance, false, A.brute_force != 0);

    for (uint32_t depth = 0; depth < A.max_depth; depth++) {
        if (depth > 0) isect = trace(S, ray.origin, ray.direction, ray.min_distance, ray.max_distance, false,
                                     A.brute_force != 0);
        cnt.rays[depth]++;
        // handleIntersections (intersection.metal:43-208)
        const bool collect = isinf_(ray.bsdf_pdf) || A.sampling_mode != 1u;
        Prng prng = ray.prng;
        ShadingContext shading;
        shading.ray_flags = ray.flags;
        shading.rnd = prng.sample3d();
        shading.wo = -ray.direction;
        bool emit_shadow = false, emit_next = false;
        ShadowRay sray{};
        Ray next{};
        if (isect.t <= 0.0f) {
            if (collect) {
                const float mis = (A.sampling_mode == 0u) || isinf_(ray.bsdf_pdf)
                                      ? 1.0f
                                      : mis_weight(ray.bsdf_pdf, lights_envmap_pdf(ctx, ray.direction));
                evaluate_environment(ctx, shading);
                add(px, mis * ray.weight * shading.material.emission);
            }
        } else {
            const PerInstance& inst = ctx.instances[isect.inst];
            uint32_t shader_index;
            shading.build(ctx, inst, isect.prim, vec2(isect.b0, isect.b1), isect.t, shader_index);
            if (inst.visibility & ray.flags) shade_surface(shader_index, ctx, shading);
            else shading.material.alpha = 0.0f;
            if (collect && mean(shading.material.emission) != 0.0f) {
                const float mis = (A.sampling_mode == 0u) || isinf_(ray.bsdf_pdf)
                                      ? 1.0f
                                      : mis_weight(ray.bsdf_pdf, lights_shape_pdf(ctx, inst, shading));
                add(px, mis * ray.weight * shading.material.emission);
            }
            vec3 sh_normal;
            if (all_zero(shading.material.normal)) {
                sh_normal = shading.true_normal;
            } else {
                vec3 geo = shading.true_normal;
                if (dot(geo, shading.wo) < 0.0f) geo *= -1.0f;
                sh_normal = valid_reflection(geo, shading.wo, shading.material.normal);
            }
            if (A.sampling_mode != 0u) {
                const LightSample nee = lights_sample(ctx, shading, prng);
                float bsdf_pdf;
                const vec3 bsdf = shading.material.evaluate(shading.wo, nee.direction, sh_normal,
                                                            shading.true_normal, bsdf_pdf);
                const vec3 contribution = nee.weight * bsdf * ray.weight;
                if (nee.casts_shadows) {
                    const float mis = A.sampling_mode == 1u || !nee.can_be_hit ? 1.0f : mis_weight(nee.pdf, bsdf_pdf);
                    const vec3 nw = mis * contribution;
                    if (all_finite(nw) && any_nonzero(nw)) {
                        emit_shadow = true;
                        sray.origin = shading.position;
                        sray.direction = nee.direction;
                        sray.min_distance = kEps;
                        sray.max_distance = nee.distance - kEps;
                        sray.weight = nw;
                    }
                } else {
                    add(px, ray.weight * contribution);
                }
            }
            const BsdfSample smp =
                shading.material.sample(shading.rnd, -ray.direction, sh_normal, shading.true_normal, ray.flags);
            const vec3 weight = ray.weight * smp.weight;
            const float mean_w = mean(weight);
            if (isfinite_(mean_w)) {
                float survival;
                bool live;
                if (depth < A.rr_depth) {
                    survival = 1.0f;
                    live = mean_w > 0.0f;
                } else {
                    survival = min_(mean_w, 1.0f);
                    live = prng.sample() < survival;
                }
                if (live) {
                    emit_next = true;
                    next.origin = shading.position;
                    next.flags = smp.flags;
                    next.depth = ray.depth + 1;
                    next.direction = smp.wi;
                    next.min_distance = kEps;
                    next.max_distance = kInf;
                    next.weight = weight * (1.0f / survival);
                    next.x = ray.x;
                    next.y = ray.y;
                    next.prng = prng;
                    next.bsdf_pdf = smp.pdf;
                }
            }
        }
        // the last depth's shadow rays are never traced (Renderer.swift:623-625)
        if (depth + 1 == A.max_depth) break;
        if (emit_shadow) {
            cnt.shadows[depth]++;
            // handleShadowRays (shadow.metal:6-29): unoccluded -> add weight
            const Hit h = trace(S, sray.origin, sray.direction, sray.min_distance, sray.max_distance, true,
                                A.brute_force != 0);
            if (h.t < 0.0f) add(px, sray.weight);
        }
        if (!emit_next) break;
        ray = next;
    }
}

void copy_light_info(rm::LightInfo& o, const rm_light_info& i) {
    o.shader_index = i.shader_index;
    o.casts_shadows = i.casts_shadows;
    o.uses_mis = i.uses_mis;
}

}  // namespace

extern "C" {

or_scene* or_scene_create(const rm_scene_desc* d, const char* material_so, char* err, size_t err_cap) {
    auto fail = [&](const std::string& m) -> or_scene* {
        if (err && err_cap) snprintf(err, err_cap, "%s", m.c_str());
        return nullptr;
    };
    if (!d || !material_so) return fail("null argument");
    or_scene* s = new or_scene();
    s->so = dlopen(material_so, RTLD_NOW | RTLD_LOCAL);
    if (!s->so) {
        const std::string m = dlerror();
        delete s;
        return fail("dlopen: " + m);
    }
    s->mat.surface = (MatFn)dlsym(s->so, "or_material_surface");
    s->mat.light = (MatFn)dlsym(s->so, "or_material_light");
    if (!s->mat.surface || !s->mat.light) {
        dlclose(s->so);
        delete s;
        return fail("material object lacks or_material_surface/or_material_light");
    }
    const size_t nv = d->num_vertices, nf = d->num_faces;
    s->vertices.assign(d->vertices, d->vertices + 3 * nv);
    s->normals.assign(d->normals, d->normals + 3 * nv);
    if (d->texcoords) s->texcoords.assign(d->texcoords, d->texcoords + 2 * nv);
    else s->texcoords.assign(2 * nv, 0.0f);
    s->indices.assign(d->indices, d->indices + 3 * nf);
    s->materials.assign(d->face_materials, d->face_materials + nf);
    s->env_res = d->envmap_resolution ? (int)d->envmap_resolution : 2048;

    // PerInstanceData (EntityBuilder.swift:97-141), shape lights (:103-114)
    uint32_t light_face_offset = 0;
    for (uint32_t k = 0; k < d->num_instances; k++) {
        const rm_instance_desc& id = d->instances[k];
        s->inst_faces.push_back(id.face_count);
        rm::PerInstance pi{};
        pi.vertex_offset = id.vertex_offset;
        pi.face_offset = id.face_offset;
        pi.visibility = id.visibility;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t v = 0; v < id.vertex_count; v++)
            for (int a = 0; a < 3; a++) {
                const float c = s->vertices[3 * (id.vertex_offset + v) + a];
                lo[a] = std::min(lo[a], c);
                hi[a] = std::max(hi[a], c);
            }
        if (id.vertex_count == 0)
            for (int a = 0; a < 3; a++) lo[a] = hi[a] = 0.0f;
        if (id.has_bounds)
            for (int a = 0; a < 3; a++) {
                lo[a] = id.bounds_min[a];
                hi[a] = id.bounds_max[a];
            }
        pi.bounds_min = vec3(lo[0], lo[1], lo[2]);
        pi.bounds_size = vec3(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
        pi.point_transform = mat4_of(id.transform);
        pi.normal_transform = inverse_transpose(pi.point_transform);
        pi.light_index = 0xffffffffu;
        if (id.emissive) {
            pi.light_index = (uint32_t)s->shape.size();
            pi.light_face_offset = light_face_offset;
            pi.light_face_count = id.face_count;
            s->light_faces.resize(light_face_offset + id.face_count);
            rm::mat3 inner;
            inner.c0 = pi.point_transform.c0.xyz();
            inner.c1 = pi.point_transform.c1.xyz();
            inner.c2 = pi.point_transform.c2.xyz();
            const float area = light_distribution(inner, &s->indices[3 * id.face_offset],
                                                  &s->vertices[3 * id.vertex_offset], id.face_count,
                                                  s->light_faces.data() + light_face_offset);
            s->shape.push_back(rm::ShapeLight{k, area});
            light_face_offset += id.face_count;
        }
        s->instances.push_back(pi);
        // world-space triangles, (instance, face) order
        for (uint32_t f = 0; f < id.face_count; f++) {
            const uint32_t* ix = &s->indices[3 * (id.face_offset + f)];
            WTri t;
            auto P = [&](uint32_t i) {
                const float* p = &s->vertices[3 * (id.vertex_offset + i)];
                return pi.point_transform.point(vec3(p[0], p[1], p[2]));
            };
            t.v0 = P(ix[0]);
            t.v1 = P(ix[1]);
            t.v2 = P(ix[2]);
            t.inst = k;
            t.prim = f;
            s->tris.push_back(t);
        }
    }
    for (uint32_t i = 0; i < d->num_area_lights; i++) {
        const rm_area_light& a = d->area_lights[i];
        rm::AreaLight L{};
        copy_light_info(L.info, a.info);
        L.t0 = rm::vec4(a.transform[0], a.transform[1], a.transform[2], a.transform[3]);
        L.t1 = rm::vec4(a.transform[4], a.transform[5], a.transform[6], a.transform[7]);
        L.t2 = rm::vec4(a.transform[8], a.transform[9], a.transform[10], a.transform[11]);
        L.color = vec3(a.color[0], a.color[1], a.color[2]);
        L.is_circular = a.is_circular;
        s->area.push_back(L);
    }
    for (uint32_t i = 0; i < d->num_point_lights; i++) {
        const rm_point_light& a = d->point_lights[i];
        rm::PointLight L{};
        copy_light_info(L.info, a.info);
        L.location = vec3(a.location[0], a.location[1], a.location[2]);
        L.radius = a.radius;
        L.color = vec3(a.color[0], a.color[1], a.color[2]);
        s->point.push_back(L);
    }
    for (uint32_t i = 0; i < d->num_sun_lights; i++) {
        const rm_sun_light& a = d->sun_lights[i];
        rm::SunLight L{};
        copy_light_info(L.info, a.info);
        L.direction = vec3(a.direction[0], a.direction[1], a.direction[2]);
        L.cos_angle = a.cos_angle;
        L.color = vec3(a.color[0], a.color[1], a.color[2]);
        s->sun.push_back(L);
    }
    for (uint32_t i = 0; i < d->num_spot_lights; i++) {
        const rm_spot_light& a = d->spot_lights[i];
        rm::SpotLight L{};
        copy_light_info(L.info, a.info);
        L.location = vec3(a.location[0], a.location[1], a.location[2]);
        L.direction = vec3(a.direction[0], a.direction[1], a.direction[2]);
        L.radius = a.radius;
        L.color = vec3(a.color[0], a.color[1], a.color[2]);
        L.spot_size = a.spot_size;
        L.spot_blend = a.spot_blend;
        s->spot.push_back(L);
    }
    s->bvh_idx.resize(s->tris.size());
    for (size_t i = 0; i < s->tris.size(); i++) s->bvh_idx[i] = (uint32_t)i;
    if (!s->tris.empty()) bvh_build(*s, 0, (uint32_t)s->tris.size());

    rm::SceneView& v = s->view;
    v.vertices = s->vertices.data();
    v.indices = s->indices.data();
    v.normals = s->normals.data();
    v.texcoords = s->texcoords.data();
    v.instances = s->instances.data();
    v.materials = s->materials.data();
    for (uint32_t i = 0; i < d->num_textures; i++) {
        const rm_texture_desc& t = d->textures[i];
        if (!t.pixels || t.width == 0 || t.height == 0 || (t.channels != 1 && t.channels != 4)) {
            if (s->so) dlclose(s->so);
            delete s;
            return fail("bad texture " + std::to_string(i));
        }
        s->tex_pixels.emplace_back(t.pixels, t.pixels + (size_t)t.width * t.height * t.channels);
        s->tex.push_back(rm::Texture{s->tex_pixels.back().data(), (int)t.width, (int)t.height, (int)t.channels});
    }
    v.textures = s->tex.data();
    v.num_textures = d->num_textures;
    v.camera.transform = mat4_of(d->camera.transform);
    v.camera.near_clip = d->camera.near_clip;
    v.camera.far_clip = d->camera.far_clip;
    v.camera.focal_length = d->camera.focal_length;
    v.camera.shift = vec2(d->camera.shift[0], d->camera.shift[1]);
    rm::LightsView& L = v.lights;
    L.num_area = (int)s->area.size();
    L.num_point = (int)s->point.size();
    L.num_sun = (int)s->sun.size();
    L.num_spot = (int)s->spot.size();
    L.num_shape = (int)s->shape.size();
    L.num_total = 1 + L.num_area + L.num_point + L.num_sun + L.num_spot + L.num_shape;
    L.light_faces = s->light_faces.data();
    L.area = s->area.data();
    L.point = s->point.data();
    L.sun = s->sun.data();
    L.spot = s->spot.data();
    L.shape = s->shape.data();
    L.world.shader_index = d->world_shader;
    L.world.resolution = s->env_res;
    return s;
}

void or_scene_destroy(or_scene* s) {
    if (!s) return;
    if (s->so) dlclose(s->so);
    delete s;
}

// envmap build / reduce / normalize (device/kernels/envmap/*.metal, LightBuilder.swift:83-167)
int or_build_envmap(or_scene* s, int threads) {
    const int res = s->env_res;
    int levels = 0;
    while ((1 << levels) < res) levels++;
    size_t mip_size = 0;
    for (int k = 0; k <= levels; k++) mip_size += (size_t)1 << (2 * k);
    s->env_mip.assign(mip_size, 0.0f);
    s->env_pdfs.assign((size_t)res * res, 0.0f);
    s->view.lights.world.pdfs = s->env_pdfs.data();
    s->view.lights.world.mipmap = s->env_mip.data();
    float* last = s->env_mip.data() + (mip_size - (size_t)res * res);
    const MatTable* mt = &s->mat;
    parallel_for((uint32_t)res * (uint32_t)res, threads, [&](uint32_t tid, int) {
        t_mat = mt;
        const uint32_t x = tid % (uint32_t)res, y = tid / (uint32_t)res;
        const uint32_t ray_index = y * (uint32_t)res + x;
        float value = 0.0f;
        for (int smp = 0; smp < 64; smp++) {
            rm::Prng prng((uint32_t)smp, ray_index);
            const vec2 j = prng.sample2d();
            const vec2 projected(((float)x + j.x) / (float)res, ((float)y + j.y) / (float)res);
            const vec3 wo = rm::warp::uniform_square_to_sphere(projected);
            rm::ShadingContext shading;
            shading.ray_flags = 0;
            shading.rnd = prng.sample3d();
            shading.wo = -wo;
            rm::evaluate_environment(s->view, shading);
            value += rm::mean(shading.material.emission);
        }
        value /= 64.0f;
        value += 1e-8f;
        const uint32_t quad = (y / 2) * ((uint32_t)res / 2) + (x / 2);
        last[4 * quad + (x & 1) + 2 * (y & 1)] = value;
        s->env_pdfs[(size_t)y * res + x] = value;
    });
    size_t off = mip_size - (size_t)res * res;
    for (int k = levels - 1; k >= 0; k--) {
        const int cr = 1 << k;
        off -= (size_t)cr * cr;
        float* level = s->env_mip.data() + off;
        const uint32_t len = (uint32_t)(cr * cr);
        for (uint32_t y = 0; y < (uint32_t)cr; y++)
            for (uint32_t x = 0; x < (uint32_t)cr; x++) {
                const uint32_t input = y * (uint32_t)cr + x;
                float sum = 0.0f;
                for (int i = 0; i < 4; ++i) sum += level[len + 4 * input + i];
                for (int i = 0; i < 4; ++i) level[len + 4 * input + i] /= sum;
                const uint32_t quad = (y / 2) * ((uint32_t)cr / 2) + (x / 2);
                level[4 * quad + (x & 1) + 2 * (y & 1)] = sum;
            }
    }
    const float total = s->env_mip[0];
    const uint32_t n = (uint32_t)res * (uint32_t)res;
    for (uint32_t i = 0; i < n; i++) s->env_pdfs[i] *= ((float)n * (1.0f / (4.0f * rm::kPi))) / total;
    return 0;
}

int or_set_envmap(or_scene* s, const float* pdfs, const float* mipmap) {
    const int res = s->env_res;
    int levels = 0;
    while ((1 << levels) < res) levels++;
    size_t mip_size = 0;
    for (int k = 0; k <= levels; k++) mip_size += (size_t)1 << (2 * k);
    s->env_mip.assign(mipmap, mipmap + mip_size);
    s->env_pdfs.assign(pdfs, pdfs + (size_t)res * res);
    s->view.lights.world.pdfs = s->env_pdfs.data();
    s->view.lights.world.mipmap = s->env_mip.data();
    return 0;
}

int or_read_envmap(or_scene* s, float* pdfs, float* mipmap) {
    if (s->env_pdfs.empty()) return -1;
    if (pdfs) std::memcpy(pdfs, s->env_pdfs.data(), s->env_pdfs.size() * 4);
    if (mipmap) std::memcpy(mipmap, s->env_mip.data(), s->env_mip.size() * 4);
    return 0;
}

int or_read_light_tables(or_scene* s, float* light_faces, float* emissive_areas) {
    if (light_faces && !s->light_faces.empty()) std::memcpy(light_faces, s->light_faces.data(), s->light_faces.size() * 4);
    if (emissive_areas)
        for (size_t i = 0; i < s->shape.size(); i++) emissive_areas[i] = s->shape[i].emissive_area;
    return 0;
}

int or_trace_closest(or_scene* s, const float* rays, uint32_t n, float* t, uint32_t* inst, uint32_t* prim, float* bary,
                     int brute_force, int threads) {
    parallel_for(n, threads, [&](uint32_t i, int) {
        const float* r = rays + 8 * (size_t)i;
        const Hit h = trace(*s, vec3(r[0], r[1], r[2]), vec3(r[4], r[5], r[6]), r[3], r[7], false, brute_force != 0);
        t[i] = h.t;
        inst[i] = h.inst;
        prim[i] = h.prim;
        bary[2 * i] = h.b0;
        bary[2 * i + 1] = h.b1;
    });
    return 0;
}

int or_trace_any(or_scene* s, const float* rays, uint32_t n, uint32_t* occluded, int brute_force, int threads) {
    parallel_for(n, threads, [&](uint32_t i, int) {
        const float* r = rays + 8 * (size_t)i;
        const Hit h = trace(*s, vec3(r[0], r[1], r[2]), vec3(r[4], r[5], r[6]), r[3], r[7], true, brute_force != 0);
        occluded[i] = h.t >= 0.0f ? 1u : 0u;
    });
    return 0;
}

static int render_impl(or_scene* s, const or_render_args* A, float* image, double* image64, or_render_result* res) {
    if (!s || !A || (!image && !image64)) return -1;
    if (s->env_pdfs.empty()) return -2;
    const uint32_t np = A->pixels ? A->num_pixels : A->width * A->height;
    int nt = A->threads > 0 ? A->threads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    std::vector<PathCounts> counts(nt);
    const MatTable* mt = &s->mat;
    rm::Lens lens{};
    std::vector<rm::LensSurface> lens_surfaces;
    if (A->lens && A->lens->num_surfaces > 0) {
        for (uint32_t i = 0; i < A->lens->num_surfaces; i++) {
            const rm_lens_surface& a = A->lens->surfaces[i];
            lens_surfaces.push_back(rm::LensSurface{a.radius, a.thickness, a.aperture, a.ior_d, a.abbe});
        }
        lens.surfaces = lens_surfaces.data();
        lens.num_surfaces = (int)A->lens->num_surfaces;
        lens.stop_index = A->lens->stop_index;
        lens.relative_stop = A->lens->relative_stop;
        lens.aperture_blades = A->lens->aperture_blades;
        lens.sensor_scale = A->lens->sensor_scale;
        lens.camera_scale = A->lens->camera_scale;
        lens.focus = A->lens->focus;
        lens.spectral = A->lens->spectral ? 1 : 0;
    }
    const auto t0 = std::chrono::steady_clock::now();
    // full images go in 8x8 pixel tiles, raster order of tiles
    const uint32_t tiles_x = (A->width + 7) / 8, tiles_y = (A->height + 7) / 8;
    const uint32_t n_iter = A->pixels ? A->num_pixels : tiles_x * tiles_y * 64;
    parallel_for(n_iter, nt, [&](uint32_t i, int th) {
        t_mat = mt;
        uint32_t x, y;
        if (A->pixels) {
            x = A->pixels[i] % A->width;
            y = A->pixels[i] / A->width;
        } else {
            const uint32_t tile = i / 64, within = i % 64;
            x = (tile % tiles_x) * 8 + within % 8;
            y = (tile / tiles_x) * 8 + within / 8;
            if (x >= A->width || y >= A->height) return;
        }
        const size_t at = 4 * ((size_t)y * A->width + x);
        for (uint32_t f = 0; f < A->frames; f++) {
            if (image) {
                render_path(*s, *A, lens, x, y, A->frame_begin + f, image + at, counts[th]);
            } else {
                // each sample summed in FP32 (its contributions in path order),
                // the samples in FP64
                float px[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                render_path(*s, *A, lens, x, y, A->frame_begin + f, px, counts[th]);
                for (int c = 0; c < 4; c++) image64[at + c] += (double)px[c];
            }
        }
    });
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (res) {
        std::memset(res, 0, sizeof(*res));
        for (auto& c : counts)
            for (int d = 0; d < 32; d++) {
                res->rays[d] += c.rays[d];
                res->shadow_rays[d] += c.shadows[d];
            }
        for (int d = 0; d < 32; d++) res->total_rays += res->rays[d] + res->shadow_rays[d];
        res->paths = (uint64_t)np * A->frames;
        res->seconds = secs;
        res->threads_used = nt;
    }
    return 0;
}

int or_render(or_scene* s, const or_render_args* A, float* image, or_render_result* res) {
    if (!image) return -1;
    return render_impl(s, A, image, nullptr, res);
}

int or_render_f64(or_scene* s, const or_render_args* A, double* image, or_render_result* res) {
    if (!image) return -1;
    return render_impl(s, A, nullptr, image, res);
}

int or_scene_set_two_level(or_scene* s, int on) {
    s->two_level = on != 0;
    s->tlas.assign(s->instances.size(), rm::TlasInstance{});
    for (size_t k = 0; k < s->instances.size(); k++)
        if (!rm::tlas_inverse(s->instances[k].point_transform, s->tlas[k].m)) return -1;
    return 0;
}

}  // extern "C"
