// Host-only structural check of the BVH8 builder: every triangle appears once
// in the leaves and every decoded (quantised) child box contains the exact
// bounds of its subtree.  Exposed through include/rm_host.h for tests.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "scene_api.h"
#include "../../../include/rm_host.h"
#include "bvh_build.h"

namespace rm {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::fmin(lo[a], b.lo[a]);
            hi[a] = std::fmax(hi[a], b.hi[a]);
        }
    }
};

Box tri_box(const BvhTri& t) {
    Box b;
    const float v[3][3] = {{t.v0x, t.v0y, t.v0z}, {t.v1x, t.v1y, t.v1z}, {t.v2x, t.v2y, t.v2z}};
    for (auto& p : v)
        for (int a = 0; a < 3; a++) {
            b.lo[a] = std::fmin(b.lo[a], p[a]);
            b.hi[a] = std::fmax(b.hi[a], p[a]);
        }
    return b;
}

struct Checker {
    const Bvh8Build& b;
    std::vector<int> seen;
    std::string err;
    int max_depth = 0;

    Box node(uint32_t idx, int depth) {
        max_depth = std::max(max_depth, depth);
        Box total;
        if (idx >= b.nodes.size()) {
            err = "node index out of range";
            return total;
        }
        const Bvh8Node& n = b.nodes[idx];
        const float p[3] = {n.px, n.py, n.pz};
        const float sc[3] = {exp_scale(n.ex), exp_scale(n.ey), exp_scale(n.ez)};
        const uint8_t* qlo[3] = {n.qlo_x, n.qlo_y, n.qlo_z};
        const uint8_t* qhi[3] = {n.qhi_x, n.qhi_y, n.qhi_z};
        int rank = 0;
        for (int c = 0; c < 8 && err.empty(); c++) {
            const uint8_t meta = n.meta[c];
            if (meta == 0) continue;
            Box sub;
            const bool internal = (n.imask >> c) & 1;
            if (internal) {
                if ((meta & 31) != 24 + c || (meta >> 5) != 1) err = "bad internal meta";
                sub = node(n.child_base + rank, depth + 1);
                rank++;
            } else {
                const uint32_t cnt_bits = meta >> 5, off = meta & 31;
                if (cnt_bits != 1 && cnt_bits != 3 && cnt_bits != 7) err = "bad leaf count";
                const uint32_t cnt = cnt_bits == 1 ? 1 : (cnt_bits == 3 ? 2 : 3);
                if (off + cnt > 24) err = "leaf offset out of range";
                for (uint32_t t = 0; t < cnt; t++) {
                    const uint32_t ti = n.tri_base + off + t;
                    if (ti >= b.tris.size()) {
                        err = "triangle index out of range";
                        break;
                    }
                    seen[ti]++;
                    sub.grow(tri_box(b.tris[ti]));
                }
            }
            for (int a = 0; a < 3; a++) {
                const float dlo = decode_q(p[a], qlo[a][c], sc[a]);
                const float dhi = decode_q(p[a], qhi[a][c], sc[a]);
                if (!(dlo <= sub.lo[a] && dhi >= sub.hi[a])) {
                    char buf[160];
                    snprintf(buf, sizeof(buf), "child box not conservative: node %u slot %d axis %d [%g,%g] vs [%g,%g]",
                             idx, c, a, dlo, dhi, sub.lo[a], sub.hi[a]);
                    err = buf;
                }
            }
            total.grow(sub);
        }
        return total;
    }
};

}  // namespace
}  // namespace rm

extern "C" int rmh_bvh8_check(const float* tris, uint32_t n, int threads, uint32_t* out_nodes, uint32_t* out_depth,
                              char* err, size_t err_cap) {
    std::vector<rm::BuildTri> in(n);
    for (uint32_t i = 0; i < n; i++) {
        const float* t = tris + 9 * (size_t)i;
        in[i].v0 = rm::vec3(t[0], t[1], t[2]);
        in[i].v1 = rm::vec3(t[3], t[4], t[5]);
        in[i].v2 = rm::vec3(t[6], t[7], t[8]);
        in[i].inst = 0;
        in[i].prim = i;
    }
    std::string e;
    rm::Bvh8Build b = rm::build_bvh8(in, threads, &e);
    auto fail = [&](const std::string& m) {
        if (err && err_cap) snprintf(err, err_cap, "%s", m.c_str());
        return -1;
    };
    if (!e.empty()) return fail(e);
    if (b.tris.size() != n) return fail("leaf triangle count differs from input");
    rm::Checker c{b, std::vector<int>(b.tris.size(), 0), "", 0};
    if (n) c.node(0, 1);
    if (!c.err.empty()) return fail(c.err);
    std::vector<int> prim_seen(n, 0);
    for (size_t i = 0; i < b.tris.size(); i++) {
        if (c.seen[i] != 1) return fail("leaf triangle referenced " + std::to_string(c.seen[i]) + " times");
        prim_seen[b.tris[i].prim]++;
    }
    for (uint32_t i = 0; i < n; i++)
        if (prim_seen[i] != 1) return fail("input triangle missing or duplicated");
    if (out_nodes) *out_nodes = (uint32_t)b.nodes.size();
    if (out_depth) *out_depth = (uint32_t)c.max_depth;
    return 0;
}

namespace {
struct HostStack2l {
    rm::Entry2l e[rm::kMaxStack2l];
    void store(int sp, rm::Entry2l v) { e[sp] = v; }
    rm::Entry2l load(int sp) const { return e[sp]; }
};
}  // namespace

extern "C" int rmh_two_level_trace(const float* vertices, uint32_t num_vertices, const uint32_t* indices,
                                   uint32_t num_faces, const rm_instance_desc* instances, uint32_t num_instances,
                                   const float* rays, uint32_t n, int any, int threads, float* t, uint32_t* inst,
                                   uint32_t* prim, float* bary, uint64_t* out_stored, char* err, size_t err_cap) {
    try {
        for (uint32_t k = 0; k < num_instances; k++) {
            const rm_instance_desc& id = instances[k];
            RM_REQUIRE((uint64_t)id.face_offset + id.face_count <= num_faces, "instance faces out of range");
            for (uint32_t f = 0; f < id.face_count; f++)
                for (int j = 0; j < 3; j++)
                    RM_REQUIRE((uint64_t)id.vertex_offset + indices[3 * ((size_t)id.face_offset + f) + j] < num_vertices,
                               "instance vertex index out of range");
        }
        const rm::TwoLevelBuild b = rm::build_two_level(std::vector<float>(vertices, vertices + 3 * (size_t)num_vertices),
                                                        std::vector<uint32_t>(indices, indices + 3 * (size_t)num_faces),
                                                        std::vector<rm_instance_desc>(instances, instances + num_instances),
                                                        threads);
        for (uint32_t i = 0; i < n; i++) {
            const float* r = rays + 8 * (size_t)i;
            HostStack2l stk;
            rm::Hit2l h;
            const rm::vec3 o(r[0], r[1], r[2]), d(r[4], r[5], r[6]);
            if (any)
                rm::traverse2l<true>(b.nodes.data(), b.tris.data(), b.recs.data(), b.root, o, d, r[3], r[7], stk, h);
            else
                rm::traverse2l<false>(b.nodes.data(), b.tris.data(), b.recs.data(), b.root, o, d, r[3], r[7], stk, h);
            t[i] = h.found ? h.t : -1.0f;
            inst[i] = h.inst;
            prim[i] = h.prim;
            bary[2 * (size_t)i] = h.u / h.det;
            bary[2 * (size_t)i + 1] = h.v / h.det;
        }
        if (out_stored) *out_stored = b.stored;
        return 0;
    } catch (const rm::HipError& e) {
        if (err && err_cap) snprintf(err, err_cap, "%s", e.msg.c_str());
        return -1;
    } catch (const std::exception& e) {
        if (err && err_cap) snprintf(err, err_cap, "%s", e.what());
        return -1;
    }
}
