// Scene construction: geometry upload, per-instance data, light tables, BVH,
// material JIT and the environment importance map.
//
// Restates the device-facing parts of host/scene/ShapeBuilder.swift:251-340,
// EntityBuilder.swift:79-173, LightBuilder.swift:83-281 and
// host/lights/distribution.m:7-39.
#include <array>
#include <thread>
#include <tuple>
#include <map>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "../device/bvh2l.h"
#include "../kernels/misc_api.h"
#include "../kernels/trace_api.h"
#include "../device/shade_kernels_abi.h"
#include "bvh_build.h"
#include "core.h"
#include "scene_api.h"

namespace rm {

static mat4 mat4_from(const float* m) {
    mat4 r;
    r.c0 = vec4(m[0], m[1], m[2], m[3]);
    r.c1 = vec4(m[4], m[5], m[6], m[7]);
    r.c2 = vec4(m[8], m[9], m[10], m[11]);
    r.c3 = vec4(m[12], m[13], m[14], m[15]);
    return r;
}

static mat3 inner3x3(const mat4& m) {
    mat3 r;
    r.c0 = m.c0.xyz();
    r.c1 = m.c1.xyz();
    r.c2 = m.c2.xyz();
    return r;
}

// inverse-transpose of the inner 3x3 = cofactor matrix / determinant, in double
mat3 normal_transform_of(const mat4& m4) {
    double m[3][3];  // m[row][col]
    const vec4 cols[3] = {m4.c0, m4.c1, m4.c2};
    for (int c = 0; c < 3; c++) {
        m[0][c] = cols[c].x;
        m[1][c] = cols[c].y;
        m[2][c] = cols[c].z;
    }
    double C[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
            C[r][c] = m[(r + 1) % 3][(c + 1) % 3] * m[(r + 2) % 3][(c + 2) % 3] -
                      m[(r + 1) % 3][(c + 2) % 3] * m[(r + 2) % 3][(c + 1) % 3];
    const double det = m[0][0] * C[0][0] + m[0][1] * C[0][1] + m[0][2] * C[0][2];
    mat3 r;
    vec3* out[3] = {&r.c0, &r.c1, &r.c2};
    for (int c = 0; c < 3; c++) *out[c] = vec3((float)(C[0][c] / det), (float)(C[1][c] / det), (float)(C[2][c] / det));
    return r;
}

// Face-area CDF of one emissive instance (host/lights/distribution.m:7-39):
// areas from the non-inverted 3x3, accumulated in double, every face counted.
float build_light_distribution(const mat3& M, const uint32_t* idx, const float* verts, uint32_t face_count,
                               float* out) {
    double acc = 0.0;
    for (uint32_t i = 0; i < face_count; i++) {
        const uint32_t a = idx[3 * i], b = idx[3 * i + 1], c = idx[3 * i + 2];
        const vec3 v0(verts[3 * a], verts[3 * a + 1], verts[3 * a + 2]);
        const vec3 v1(verts[3 * b], verts[3 * b + 1], verts[3 * b + 2]);
        const vec3 v2(verts[3 * c], verts[3 * c + 1], verts[3 * c + 2]);
        const vec3 tn = cross(M * (v2 - v0), M * (v2 - v1));
        const float area = length(tn) / 2.0f;
        acc += (double)area;
        out[i] = (float)acc;
    }
    for (uint32_t i = 0; i < face_count; i++) out[i] = (float)((double)out[i] / acc);
    return (float)acc;
}

static LightInfo info_of(const rm_light_info& i) { return LightInfo{i.shader_index, i.casts_shadows, i.uses_mis}; }

void scene_create(rm_device* dev, const rm_scene_desc* d, rm_scene* s) {
    RM_REQUIRE(d != nullptr, "null scene description");
    RM_REQUIRE(d->num_vertices == 0 || (d->vertices && d->normals), "vertices/normals missing");
    RM_REQUIRE(d->num_faces == 0 || (d->indices && d->face_materials), "indices/materials missing");
    RM_REQUIRE(d->num_instances == 0 || d->instances, "instances missing");
    s->dev = dev;
    const size_t nv = d->num_vertices, nf = d->num_faces;
    s->vertices.assign(d->vertices, d->vertices + 3 * nv);
    s->normals.assign(d->normals, d->normals + 3 * nv);
    if (d->texcoords) s->texcoords.assign(d->texcoords, d->texcoords + 2 * nv);
    else s->texcoords.assign(2 * nv, 0.0f);
    s->indices.assign(d->indices, d->indices + 3 * nf);
    s->materials.assign(d->face_materials, d->face_materials + nf);
    s->inst_desc.assign(d->instances, d->instances + d->num_instances);
    s->world_shader = d->world_shader;
    s->envmap_res = d->envmap_resolution ? (int)d->envmap_resolution : 2048;
    RM_REQUIRE((s->envmap_res & (s->envmap_res - 1)) == 0, "envmap resolution must be a power of two");

    // per-instance data and shape lights (EntityBuilder.swift:97-141)
    uint32_t light_face_offset = 0;
    for (uint32_t k = 0; k < d->num_instances; k++) {
        const rm_instance_desc& id = d->instances[k];
        RM_REQUIRE((size_t)id.vertex_offset + id.vertex_count <= nv, "instance vertex range out of bounds");
        RM_REQUIRE((size_t)id.face_offset + id.face_count <= nf, "instance face range out of bounds");
        for (uint32_t f = 0; f < id.face_count; f++)
            for (int c = 0; c < 3; c++)
                RM_REQUIRE(s->indices[3 * (id.face_offset + f) + c] < id.vertex_count, "face index out of range");
        PerInstance pi{};
        pi.vertex_offset = id.vertex_offset;
        pi.face_offset = id.face_offset;
        pi.visibility = id.visibility;
        vec3 bmin(INFINITY), bmax(-INFINITY);
        for (uint32_t v = 0; v < id.vertex_count; v++) {
            const float* p = &s->vertices[3 * (id.vertex_offset + v)];
            bmin = vmin(bmin, vec3(p[0], p[1], p[2]));
            bmax = vmax(bmax, vec3(p[0], p[1], p[2]));
        }
        if (id.vertex_count == 0) bmin = bmax = vec3(0.0f);
        if (id.has_bounds) {
            bmin = vec3(id.bounds_min[0], id.bounds_min[1], id.bounds_min[2]);
            bmax = vec3(id.bounds_max[0], id.bounds_max[1], id.bounds_max[2]);
        }
        pi.bounds_min = bmin;
        pi.bounds_size = bmax - bmin;
        pi.point_transform = mat4_from(id.transform);
        pi.normal_transform = normal_transform_of(pi.point_transform);
        pi.light_index = 0xffffffffu;
        if (id.emissive) {
            pi.light_index = (uint32_t)s->shape.size();
            pi.light_face_offset = light_face_offset;
            pi.light_face_count = id.face_count;
            s->light_faces.resize(light_face_offset + id.face_count);
            const float area = build_light_distribution(inner3x3(pi.point_transform), &s->indices[3 * id.face_offset],
                                                        &s->vertices[3 * id.vertex_offset], id.face_count,
                                                        s->light_faces.data() + light_face_offset);
            s->shape.push_back(ShapeLight{k, area});
            light_face_offset += id.face_count;
        }
        s->instances.push_back(pi);
    }
    for (uint32_t i = 0; i < d->num_area_lights; i++) {
        const rm_area_light& a = d->area_lights[i];
        AreaLight L{};
        L.info = info_of(a.info);
        L.t0 = vec4(a.transform[0], a.transform[1], a.transform[2], a.transform[3]);
        L.t1 = vec4(a.transform[4], a.transform[5], a.transform[6], a.transform[7]);
        L.t2 = vec4(a.transform[8], a.transform[9], a.transform[10], a.transform[11]);
        L.color = vec3(a.color[0], a.color[1], a.color[2]);
        L.is_circular = a.is_circular;
        s->area.push_back(L);
    }
    for (uint32_t i = 0; i < d->num_point_lights; i++) {
        const rm_point_light& a = d->point_lights[i];
        s->point.push_back(PointLight{info_of(a.info), vec3(a.location[0], a.location[1], a.location[2]), a.radius,
                                      vec3(a.color[0], a.color[1], a.color[2])});
    }
    for (uint32_t i = 0; i < d->num_sun_lights; i++) {
        const rm_sun_light& a = d->sun_lights[i];
        s->sun.push_back(SunLight{info_of(a.info), vec3(a.direction[0], a.direction[1], a.direction[2]), a.cos_angle,
                                  vec3(a.color[0], a.color[1], a.color[2])});
    }
    for (uint32_t i = 0; i < d->num_spot_lights; i++) {
        const rm_spot_light& a = d->spot_lights[i];
        s->spot.push_back(SpotLight{info_of(a.info), vec3(a.location[0], a.location[1], a.location[2]),
                                    vec3(a.direction[0], a.direction[1], a.direction[2]), a.radius,
                                    vec3(a.color[0], a.color[1], a.color[2]), a.spot_size, a.spot_blend});
    }

    hipStream_t st = dev->stream;
    s->d_vertices.upload(s->vertices.data(), s->vertices.size() * 4, st);
    s->d_normals.upload(s->normals.data(), s->normals.size() * 4, st);
    s->d_texcoords.upload(s->texcoords.data(), s->texcoords.size() * 4, st);
    s->d_indices.upload(s->indices.data(), s->indices.size() * 4, st);
    s->d_materials.upload(s->materials.data(), s->materials.size() * 2, st);
    s->d_instances.upload(s->instances.data(), s->instances.size() * sizeof(PerInstance), st);
    s->d_area.upload(s->area.data(), s->area.size() * sizeof(AreaLight), st);
    s->d_point.upload(s->point.data(), s->point.size() * sizeof(PointLight), st);
    s->d_sun.upload(s->sun.data(), s->sun.size() * sizeof(SunLight), st);
    s->d_spot.upload(s->spot.data(), s->spot.size() * sizeof(SpotLight), st);
    s->d_shape.upload(s->shape.data(), s->shape.size() * sizeof(ShapeLight), st);
    s->d_light_faces.upload(s->light_faces.data(), s->light_faces.size() * 4, st);
    // texture slots of the material library
    RM_REQUIRE(d->num_textures == 0 || d->textures, "textures missing");
    std::vector<Texture> tex_table;
    for (uint32_t i = 0; i < d->num_textures; i++) {
        const rm_texture_desc& t = d->textures[i];
        RM_REQUIRE(t.width > 0 && t.height > 0 && t.pixels, "texture " + std::to_string(i) + " is empty");
        RM_REQUIRE(t.channels == 1 || t.channels == 4, "texture channels must be 1 or 4");
        RM_REQUIRE((uint64_t)t.width * t.height * t.channels < (1ull << 31), "texture too large");
        s->d_tex.emplace_back(new DevBuf());
        s->d_tex.back()->upload(t.pixels, (size_t)t.width * t.height * t.channels * 4, st);
        tex_table.push_back(Texture{s->d_tex.back()->as<float>(), (int)t.width, (int)t.height, (int)t.channels});
    }
    s->d_tex_table.upload(tex_table.data(), tex_table.size() * sizeof(Texture), st);
    RM_HIP(hipStreamSynchronize(st));

    SceneView& v = s->view;
    v.vertices = s->d_vertices.as<float>();
    v.indices = s->d_indices.as<uint32_t>();
    v.normals = s->d_normals.as<float>();
    v.texcoords = s->d_texcoords.as<float>();
    v.instances = s->d_instances.as<PerInstance>();
    v.materials = s->d_materials.as<uint16_t>();
    v.textures = s->d_tex_table.as<Texture>();
    v.num_textures = d->num_textures;
    LightsView& L = v.lights;
    L.num_area = (int)s->area.size();
    L.num_point = (int)s->point.size();
    L.num_sun = (int)s->sun.size();
    L.num_spot = (int)s->spot.size();
    L.num_shape = (int)s->shape.size();
    L.num_total = 1 + L.num_area + L.num_point + L.num_sun + L.num_spot + L.num_shape;
    L.light_faces = s->d_light_faces.as<float>();
    L.area = s->d_area.as<AreaLight>();
    L.point = s->d_point.as<PointLight>();
    L.sun = s->d_sun.as<SunLight>();
    L.spot = s->d_spot.as<SpotLight>();
    L.shape = s->d_shape.as<ShapeLight>();
    L.world.shader_index = s->world_shader;
    L.world.resolution = s->envmap_res;
    L.world.pdfs = nullptr;
    L.world.mipmap = nullptr;
    scene_set_camera(s, &d->camera);
}

void scene_set_camera(rm_scene* s, const rm_camera_desc* c) {
    RM_REQUIRE(c != nullptr, "null camera");
    Camera& cam = s->view.camera;
    cam.transform = mat4_from(c->transform);
    cam.near_clip = c->near_clip;
    cam.far_clip = c->far_clip;
    cam.focal_length = c->focal_length;
    cam.shift = vec2(c->shift[0], c->shift[1]);
}

// World-space triangles in the order (instance, face); vertices transformed
// with mat4::point (fixed association order).
std::vector<BuildTri> world_triangles(const rm_scene* s) {
    std::vector<BuildTri> tris;
    size_t total = 0;
    for (auto& id : s->inst_desc) total += id.face_count;
    tris.reserve(total);
    for (uint32_t k = 0; k < s->inst_desc.size(); k++) {
        const rm_instance_desc& id = s->inst_desc[k];
        const mat4& M = s->instances[k].point_transform;
        for (uint32_t f = 0; f < id.face_count; f++) {
            const uint32_t* ix = &s->indices[3 * (id.face_offset + f)];
            BuildTri t;
            const float* p0 = &s->vertices[3 * (id.vertex_offset + ix[0])];
            const float* p1 = &s->vertices[3 * (id.vertex_offset + ix[1])];
            const float* p2 = &s->vertices[3 * (id.vertex_offset + ix[2])];
            t.v0 = M.point(vec3(p0[0], p0[1], p0[2]));
            t.v1 = M.point(vec3(p1[0], p1[1], p1[2]));
            t.v2 = M.point(vec3(p2[0], p2[1], p2[2]));
            t.inst = k;
            t.prim = f;
            tris.push_back(t);
        }
    }
    return tris;
}

void scene_build_accel(rm_scene* s, int threads) {
    std::string err;
    const std::vector<BuildTri> tris = world_triangles(s);
    Bvh8Build b = build_bvh8(tris, threads, &err);
    RM_REQUIRE(err.empty(), "BVH build failed: " + err);
    hipStream_t st = s->dev->stream;
    s->d_nodes.upload(b.nodes.data(), b.nodes.size() * sizeof(Bvh8Node), st);
    s->d_tris.upload(b.tris.data(), b.tris.size() * sizeof(BvhTri), st);
    // spill rows of the traversal stack for the largest persistent grid
    const int bpc = std::max(s->dev->trace_blocks_per_cu, s->dev->any_blocks_per_cu);
    s->spill_threads = (size_t)s->dev->props.multiProcessorCount * (size_t)bpc * kTraceBlock;
    s->d_spill.alloc(s->spill_threads * (size_t)(kMaxTraceStack - kLdsStack) * sizeof(uint2));
    RM_HIP(hipStreamSynchronize(st));
    s->accel_built = true;
    s->two_level = false;
    s->root2l = 0;
    s->d_tlas.release();
    s->stats.triangles = tris.size();
    s->stats.stored_triangles = tris.size();
    s->stats.two_level = 0;
    s->stats.bvh_nodes = b.nodes.size();
    s->stats.bvh_depth = (uint32_t)b.depth;
    s->stats.bvh_build_seconds = b.build_seconds;
}

// Two-level form: a BVH8 per shape (distinct vertex and face range) over its
// object-space triangles, then one over the instances' world boxes with one
// instance per leaf, whose leaf records hold the world-to-object rows.  The
// world box of an instance is its shape's object box mapped in double and
// widened by 1e-5 of its largest coordinate, so that it holds every hit the
// object-space tests can report for a ray through it.
TwoLevelBuild build_two_level(const std::vector<float>& vertices, const std::vector<uint32_t>& indices,
                              const std::vector<rm_instance_desc>& insts, int threads) {
    const auto t0 = std::chrono::steady_clock::now();
    TwoLevelBuild out;
    std::string err;
    const size_t ni = insts.size();
    struct Shape {
        uint32_t vo, fo, fc, root;
    };
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, uint32_t> shape_id;
    std::vector<Shape> shapes;
    std::vector<uint32_t> inst_shape(ni);
    for (size_t k = 0; k < ni; k++) {
        const rm_instance_desc& id = insts[k];
        const auto key = std::make_tuple(id.vertex_offset, id.face_offset, id.face_count);
        auto it = shape_id.find(key);
        if (it == shape_id.end()) {
            it = shape_id.emplace(key, (uint32_t)shapes.size()).first;
            shapes.push_back(Shape{id.vertex_offset, id.face_offset, id.face_count, 0});
        }
        inst_shape[k] = it->second;
    }
    int depth = 0;
    for (Shape& sh : shapes) {
        std::vector<BuildTri> bt(sh.fc);
        for (uint32_t f = 0; f < sh.fc; f++) {
            const uint32_t* ix = &indices[3 * ((size_t)sh.fo + f)];
            vec3* v[3] = {&bt[f].v0, &bt[f].v1, &bt[f].v2};
            for (int j = 0; j < 3; j++) {
                const float* p = &vertices[3 * ((size_t)sh.vo + ix[j])];
                *v[j] = vec3(p[0], p[1], p[2]);
            }
            bt[f].inst = 0;
            bt[f].prim = f;
        }
        Bvh8Build b = build_bvh8(bt, threads, &err, 3);
        RM_REQUIRE(err.empty(), "BVH build failed: " + err);
        const uint32_t nbase = (uint32_t)out.nodes.size(), tbase = (uint32_t)out.tris.size();
        for (Bvh8Node& n : b.nodes) {
            n.child_base += nbase;
            n.tri_base += tbase;
        }
        sh.root = nbase;
        out.nodes.insert(out.nodes.end(), b.nodes.begin(), b.nodes.end());
        out.tris.insert(out.tris.end(), b.tris.begin(), b.tris.end());
        depth = std::max(depth, b.depth);
        out.stored += sh.fc;
    }
    // top level: each instance with geometry as the box (lo, hi, hi).  The
    // box bounds the shape's vertices mapped to the world in double (tighter
    // than the mapped object box: a turned shape's box grows by up to sqrt 2
    // per axis), then padded and rounded outward.
    std::vector<std::vector<uint32_t>> shape_verts(shapes.size());
    for (size_t si = 0; si < shapes.size(); si++) {
        const Shape& sh = shapes[si];
        std::vector<uint32_t>& vs = shape_verts[si];
        vs.assign(indices.begin() + 3 * (size_t)sh.fo, indices.begin() + 3 * ((size_t)sh.fo + sh.fc));
        std::sort(vs.begin(), vs.end());
        vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
    }
    std::vector<mat4> xf(ni);
    std::vector<std::array<double, 6>> wbox(ni);
    for (size_t k = 0; k < ni; k++) xf[k] = mat4_from(insts[k].transform);
    const int nthreads = std::max(1, std::min<int>(threads > 0 ? threads : (int)std::thread::hardware_concurrency(), 64));
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) {
        pool.emplace_back([&, t] {
            for (size_t k = (size_t)t; k < ni; k += (size_t)nthreads) {
                const Shape& sh = shapes[inst_shape[k]];
                const mat4& M = xf[k];
                const double m[3][4] = {{M.c0.x, M.c1.x, M.c2.x, M.c3.x}, {M.c0.y, M.c1.y, M.c2.y, M.c3.y},
                                        {M.c0.z, M.c1.z, M.c2.z, M.c3.z}};
                std::array<double, 6> b = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
                for (const uint32_t v : shape_verts[inst_shape[k]]) {
                    const float* p = &vertices[3 * ((size_t)sh.vo + v)];
                    for (int a = 0; a < 3; a++) {
                        const double w = m[a][0] * p[0] + m[a][1] * p[1] + m[a][2] * p[2] + m[a][3];
                        b[a] = std::min(b[a], w);
                        b[3 + a] = std::max(b[3 + a], w);
                    }
                }
                wbox[k] = b;
            }
        });
    }
    for (std::thread& th : pool) th.join();
    std::vector<BuildTri> boxes;
    for (size_t k = 0; k < ni; k++) {
        const Shape& sh = shapes[inst_shape[k]];
        out.total += sh.fc;
        if (sh.fc == 0) continue;
        const double lo[3] = {wbox[k][0], wbox[k][1], wbox[k][2]};
        const double hi[3] = {wbox[k][3], wbox[k][4], wbox[k][5]};
        double mag = 0.0;
        for (int a = 0; a < 3; a++) mag = std::max(mag, std::max(std::fabs(lo[a]), std::fabs(hi[a])));
        const double pad = 1e-5 * mag + 1e-30;
        BuildTri t;
        float flo[3], fhi[3];
        for (int a = 0; a < 3; a++) {
            flo[a] = (float)(lo[a] - pad);
            fhi[a] = (float)(hi[a] + pad);
            if ((double)flo[a] > lo[a] - pad) flo[a] = std::nextafter(flo[a], -INFINITY);
            if ((double)fhi[a] < hi[a] + pad) fhi[a] = std::nextafter(fhi[a], INFINITY);
        }
        t.v0 = vec3(flo[0], flo[1], flo[2]);
        t.v1 = vec3(fhi[0], fhi[1], fhi[2]);
        t.v2 = t.v1;
        t.inst = (uint32_t)k;
        t.prim = 0;
        boxes.push_back(t);
    }
    Bvh8Build top = build_bvh8(boxes, threads, &err, 1);
    RM_REQUIRE(err.empty(), "top-level BVH build failed: " + err);
    // one stack entry per level of both trees and the instance's marker
    RM_REQUIRE(depth + top.depth + 2 < kMaxTraceStack, "two-level tree too deep for the traversal stack");
    const uint32_t nbase = (uint32_t)out.nodes.size();
    for (Bvh8Node& n : top.nodes) n.child_base += nbase;
    out.nodes.insert(out.nodes.end(), top.nodes.begin(), top.nodes.end());
    out.recs.resize(top.tris.size());
    for (size_t i = 0; i < top.tris.size(); i++) {
        const uint32_t k = top.tris[i].inst;
        TlasInstance& r = out.recs[i];
        std::memset(&r, 0, sizeof(r));
        RM_REQUIRE(tlas_inverse(xf[k], r.m), "instance " + std::to_string(k) + " has a singular transform");
        r.blas_root = shapes[inst_shape[k]].root;
        r.inst = k;
    }
    // valid pointers for scenes without geometry
    if (out.recs.empty()) out.recs.resize(1);
    if (out.tris.empty()) out.tris.resize(1);
    out.root = nbase;
    out.depth = depth + top.depth;
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

void scene_build_accel_two_level(rm_scene* s, int threads) {
    const TwoLevelBuild b = build_two_level(s->vertices, s->indices, s->inst_desc, threads);
    hipStream_t st = s->dev->stream;
    s->d_nodes.upload(b.nodes.data(), b.nodes.size() * sizeof(Bvh8Node), st);
    s->d_tris.upload(b.tris.data(), b.tris.size() * sizeof(BvhTri), st);
    s->d_tlas.upload(b.recs.data(), b.recs.size() * sizeof(TlasInstance), st);
    const int bpc = std::max(s->dev->trace_blocks_per_cu, s->dev->any_blocks_per_cu);
    s->spill_threads = (size_t)s->dev->props.multiProcessorCount * (size_t)bpc * kTraceBlock;
    s->d_spill.alloc(s->spill_threads * (size_t)(kMaxTraceStack - kLdsStack) * sizeof(uint2));
    RM_HIP(hipStreamSynchronize(st));
    s->accel_built = true;
    s->two_level = true;
    s->root2l = b.root;
    s->stats.triangles = b.total;
    s->stats.stored_triangles = b.stored;
    s->stats.two_level = 1;
    s->stats.bvh_nodes = b.nodes.size();
    s->stats.bvh_depth = (uint32_t)b.depth;
    s->stats.bvh_build_seconds = b.seconds;
}

// Compiles (or takes from the code-object cache) and loads one module.
static void load_module(JitModule& j, const std::string& src, const char* options, std::string& log) {
    std::vector<char> code;
    bool hit = false;
    if (!jit_compile(src, options ? options : "", code, log, hit))
        throw HipError{RM_ERROR_COMPILE, "material compilation failed:\n" + log};
    if (j.module) {
        (void)hipModuleUnload(j.module);
        j.module = nullptr;
    }
    RM_HIP(hipModuleLoadData(&j.module, code.data()));
    RM_HIP(hipModuleGetFunction(&j.shade, j.module, "rm_shade_kernel"));
    RM_HIP(hipModuleGetFunction(&j.envmap_build, j.module, "rm_envmap_build_kernel"));
    auto optional = [&](hipFunction_t& f, const char* name) {
        if (hipModuleGetFunction(&f, j.module, name) != hipSuccess) {
            (void)hipGetLastError();
            f = nullptr;
        }
    };
    optional(j.shade_batch, "rm_shade_batch_kernel");
    optional(j.shade_cache, "rm_shade_cache_kernel");
}

void scene_cache_module(rm_scene* s) {
    if (s->jit_cache.shade_cache) return;
    RM_REQUIRE(s->materials_built, "materials must be compiled before a cache is attached");
    std::string log;
    load_module(s->jit_cache, s->material_source, "-DRM_CACHE_KERNEL", log);
    RM_REQUIRE(s->jit_cache.shade_cache != nullptr, "cache shading kernel missing from the module");
}

void scene_compile_materials(rm_scene* s, const char* src, size_t len, const char* options, std::string& log) {
    RM_REQUIRE(src != nullptr, "null material source");
    const auto t0 = std::chrono::steady_clock::now();
    load_module(s->jit, std::string(src, len), options, log);
    s->material_source.assign(src, len);
    if (s->jit_cache.module) {
        (void)hipModuleUnload(s->jit_cache.module);
        s->jit_cache.module = nullptr;
        s->jit_cache.shade_cache = nullptr;
    }
    s->materials_built = true;
    s->stats.jit_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void scene_build_envmap(rm_scene* s) {
    RM_REQUIRE(s->materials_built, "materials must be compiled before the envmap");
    const auto t0 = std::chrono::steady_clock::now();
    const int res = s->envmap_res;
    int levels = 0;
    while ((1 << levels) < res) levels++;
    size_t mip_size = 0;
    for (int k = 0; k <= levels; k++) mip_size += (size_t)1 << (2 * k);
    s->d_env_mip.alloc(mip_size * 4);
    s->d_env_pdfs.alloc((size_t)res * res * 4);
    float* mip = s->d_env_mip.as<float>();
    float* pdfs = s->d_env_pdfs.as<float>();
    s->view.lights.world.pdfs = pdfs;
    s->view.lights.world.mipmap = mip;
    hipStream_t st = s->dev->stream;
    const size_t last = mip_size - (size_t)res * res;
    EnvmapArgs a;
    a.ctx = s->view;
    a.mipmap_last = mip + last;
    a.pdfs = pdfs;
    a.resolution = res;
    void* args[] = {&a};
    const uint32_t n = (uint32_t)res * (uint32_t)res;
    RM_HIP(hipModuleLaunchKernel(s->jit.envmap_build, (n + 255) / 256, 1, 1, 256, 1, 1, 0, st, args, nullptr));
    size_t off = last;
    for (int k = levels - 1; k >= 0; k--) {
        off -= (size_t)1 << (2 * k);
        RM_HIP(launch_envmap_reduce(mip + off, 1 << k, st));
    }
    RM_HIP(launch_envmap_normalize(mip, pdfs, n, st));
    RM_HIP(hipStreamSynchronize(st));
    s->envmap_built = true;
    s->stats.envmap_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------------------
// Raw batch traversal

static void upload_rays_soa(const float* rays, uint32_t n, DevBuf& buf, hipStream_t st, float** f) {
    std::vector<float> soa((size_t)n * 8);
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 8; k++) soa[(size_t)k * n + i] = rays[(size_t)8 * i + k];
    buf.upload(soa.data(), soa.size() * 4, st);
    for (int k = 0; k < 8; k++) f[k] = buf.as<float>() + (size_t)k * n;
}

void trace_closest_raw(rm_scene* s, const float* rays, uint32_t n, float* t, uint32_t* inst, uint32_t* prim,
                       float* bary) {
    RM_REQUIRE(s->accel_built, "acceleration structure not built");
    if (n == 0) return;
    hipStream_t st = s->dev->stream;
    DevBuf in, out, cnt;
    float* f[8];
    upload_rays_soa(rays, n, in, st, f);
    out.alloc((size_t)n * 5 * 4);
    cnt.alloc(sizeof(Counters));
    RM_HIP(hipMemsetAsync(cnt.ptr, 0, sizeof(Counters), st));
    ClosestArgs a{};
    a.nodes = s->d_nodes.as<Bvh8Node>();
    a.tris = s->d_tris.as<BvhTri>();
    a.tlas = s->two_level ? s->d_tlas.as<TlasInstance>() : nullptr;
    a.root2l = s->root2l;
    a.count = n;
    a.fetch = &cnt.as<Counters>()->it[0].fetch_closest;
    a.paths.ox = f[0]; a.paths.oy = f[1]; a.paths.oz = f[2];
    a.paths.dx = f[4]; a.paths.dy = f[5]; a.paths.dz = f[6];
    a.tmin_arr = f[3];
    a.tmax_arr = f[7];
    float* o = out.as<float>();
    a.hits.t = o;
    a.hits.inst = reinterpret_cast<uint32_t*>(o + n);
    a.hits.prim = reinterpret_cast<uint32_t*>(o + 2 * (size_t)n);
    a.hits.b0 = o + 3 * (size_t)n;
    a.hits.b1 = o + 4 * (size_t)n;
    a.spill = s->d_spill.as<uint2>();
    a.counters = cnt.as<Counters>();
    const int blocks = s->dev->props.multiProcessorCount * s->dev->trace_blocks_per_cu;
    RM_HIP(launch_trace_closest(a, blocks, false, st));
    std::vector<float> host((size_t)n * 5);
    RM_HIP(hipMemcpyAsync(host.data(), out.ptr, host.size() * 4, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < n; i++) {
        if (t) t[i] = host[i];
        if (inst) std::memcpy(&inst[i], &host[n + i], 4);
        if (prim) std::memcpy(&prim[i], &host[2 * (size_t)n + i], 4);
        if (bary) {
            bary[2 * i] = host[3 * (size_t)n + i];
            bary[2 * i + 1] = host[4 * (size_t)n + i];
        }
    }
}

void trace_any_raw(rm_scene* s, const float* rays, uint32_t n, uint32_t* occluded) {
    RM_REQUIRE(s->accel_built, "acceleration structure not built");
    if (n == 0) return;
    hipStream_t st = s->dev->stream;
    DevBuf in, out, cnt;
    float* f[8];
    upload_rays_soa(rays, n, in, st, f);
    out.alloc((size_t)n * 4);
    cnt.alloc(sizeof(Counters));
    RM_HIP(hipMemsetAsync(cnt.ptr, 0, sizeof(Counters), st));
    AnyArgs a{};
    a.nodes = s->d_nodes.as<Bvh8Node>();
    a.tris = s->d_tris.as<BvhTri>();
    a.tlas = s->two_level ? s->d_tlas.as<TlasInstance>() : nullptr;
    a.root2l = s->root2l;
    a.count = n;
    a.fetch = &cnt.as<Counters>()->it[0].fetch_any;
    a.shadows.ox = f[0]; a.shadows.oy = f[1]; a.shadows.oz = f[2];
    a.shadows.dx = f[4]; a.shadows.dy = f[5]; a.shadows.dz = f[6];
    a.tmin_arr = f[3];
    a.shadows.tmax = f[7];
    a.image = nullptr;
    a.occluded = out.as<uint32_t>();
    a.spill = s->d_spill.as<uint2>();
    a.counters = cnt.as<Counters>();
    const int blocks = s->dev->props.multiProcessorCount * s->dev->any_blocks_per_cu;
    RM_HIP(launch_trace_any(a, blocks, false, st));
    RM_HIP(hipMemcpyAsync(occluded, out.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    RM_HIP(hipStreamSynchronize(st));
}

}  // namespace rm
