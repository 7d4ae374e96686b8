// Two-level BVH8 (the reference's instance acceleration structure over
// per-shape primitive structures, host/scene/EntityBuilder.swift:146-165).
//
// One node array holds every shape's bottom-level tree (object-space
// triangles, BvhTri with the shape-local primitive id) and the top-level
// tree over the instances' world boxes, built with one instance per leaf.
// A top-level leaf's record index is its TlasInstance: the world-to-object
// transform, the shape's root node and the instance id.
//
// A ray enters an instance by the affine map of its origin and direction
// (the direction is not renormalised, so t is the same parameter on both
// sides) and tests the shape's triangles in object space.  Hits are the
// closest over all instances, ties to the smaller (instance, primitive) as
// in the flat tree.
//
// The traversal keeps one stack entry per tree level, as the flat kernels
// do: a node group (child base, remaining child hits in visit order << 24 |
// internal mask).  Entering an instance pushes the top-level group and a
// marker entry holding the remaining instances of the leaf set; popping the
// marker restores the world ray.  traverse2l below is the one-ray form that
// the host check of the build runs (rmh_two_level_trace); the GPU kernels
// take the same entry and exit steps inside their per-lane step machine
// (RayState::enter and the marker pop in kernels/trace.hip), with the
// postponed leaves and lane refill of the flat kernels.
#pragma once
#include "bvh.h"
#include "isect.h"

namespace rm {

struct alignas(16) TlasInstance {
    float m[12];          // world to object: rows of a 3x4 matrix
    uint32_t blas_root;   // node index of the shape's tree
    uint32_t inst;        // instance id
    uint32_t pad0, pad1;
};
static_assert(sizeof(TlasInstance) == 64, "TlasInstance must be 64 bytes");

// A stack entry: a node group, or (kMarker2l set) the instances still to
// enter of a top-level leaf set: base = record base, bits = their mask.
struct Entry2l {
    uint32_t base, bits;
};
constexpr uint32_t kMarker2l = 1u << 23;
constexpr int kMaxStack2l = 64;

// World-to-object rows of an instance: the inverse of the affine part of its
// object-to-world transform M (columns c0..c3), in double, rounded once.
// Returns false for a singular M.
RM_HD bool tlas_inverse(const mat4& M, float out[12]) {
    const double a00 = M.c0.x, a01 = M.c1.x, a02 = M.c2.x, t0 = M.c3.x;
    const double a10 = M.c0.y, a11 = M.c1.y, a12 = M.c2.y, t1 = M.c3.y;
    const double a20 = M.c0.z, a21 = M.c1.z, a22 = M.c2.z, t2 = M.c3.z;
    const double c00 = a11 * a22 - a12 * a21, c01 = a02 * a21 - a01 * a22, c02 = a01 * a12 - a02 * a11;
    const double c10 = a12 * a20 - a10 * a22, c11 = a00 * a22 - a02 * a20, c12 = a02 * a10 - a00 * a12;
    const double c20 = a10 * a21 - a11 * a20, c21 = a01 * a20 - a00 * a21, c22 = a00 * a11 - a01 * a10;
    const double det = a00 * c00 + a01 * c10 + a02 * c20;
    if (!(det != 0.0)) return false;
    const double r = 1.0 / det;
    const double m[3][3] = {{c00 * r, c01 * r, c02 * r}, {c10 * r, c11 * r, c12 * r}, {c20 * r, c21 * r, c22 * r}};
    const double t[3] = {t0, t1, t2};
    for (int i = 0; i < 3; i++) {
        out[4 * i + 0] = (float)m[i][0];
        out[4 * i + 1] = (float)m[i][1];
        out[4 * i + 2] = (float)m[i][2];
        out[4 * i + 3] = (float)(-(m[i][0] * t[0] + m[i][1] * t[1] + m[i][2] * t[2]));
    }
    return true;
}

RM_HD vec3 tlas_point(const TlasInstance& t, vec3 p) {
    return vec3(((t.m[0] * p.x + t.m[1] * p.y) + t.m[2] * p.z) + t.m[3],
                ((t.m[4] * p.x + t.m[5] * p.y) + t.m[6] * p.z) + t.m[7],
                ((t.m[8] * p.x + t.m[9] * p.y) + t.m[10] * p.z) + t.m[11]);
}
RM_HD vec3 tlas_vector(const TlasInstance& t, vec3 v) {
    return vec3((t.m[0] * v.x + t.m[1] * v.y) + t.m[2] * v.z, (t.m[4] * v.x + t.m[5] * v.y) + t.m[6] * v.z,
                (t.m[8] * v.x + t.m[9] * v.y) + t.m[10] * v.z);
}

// Child c of node n: decoded box (conservative: the builder checks it).
RM_HD void child_box(const Bvh8Node& n, int c, vec3& lo, vec3& hi) {
    const float sx = exp_scale(n.ex), sy = exp_scale(n.ey), sz = exp_scale(n.ez);
    lo = vec3(decode_q(n.px, n.qlo_x[c], sx), decode_q(n.py, n.qlo_y[c], sy), decode_q(n.pz, n.qlo_z[c], sz));
    hi = vec3(decode_q(n.px, n.qhi_x[c], sx), decode_q(n.py, n.qhi_y[c], sy), decode_q(n.pz, n.qhi_z[c], sz));
}

struct Hit2l {
    float t;
    uint32_t inst, prim;
    float u, v, det;  // scaled barycentrics of the closest hit (v0, v1 weigh u/det, v/det)
    bool found;
};

// Closest (ANY = false) or any hit in (tmin, tmax) of the world ray (o, d).
// Stack: store(int sp, Entry2l) / load(int sp), kMaxStack2l entries.
template <bool ANY, class Stack>
RM_HD bool traverse2l(const Bvh8Node* nodes, const BvhTri* tris, const TlasInstance* insts, uint32_t root, vec3 wo,
                      vec3 wd, float tmin, float tmax, Stack& stk, Hit2l& hit) {
    vec3 o = wo, d = wd;
    vec3 inv = safe_inv_dir(d);
    RayShear sh = make_shear(d);
    int oct = dir_octant(d);
    bool top = true;          // in the top-level tree
    uint32_t inst = 0;        // the instance entered
    hit.t = tmax;
    hit.inst = 0xffffffffu;
    hit.prim = 0xffffffffu;
    hit.u = 0.0f;
    hit.v = 0.0f;
    hit.det = 1.0f;
    hit.found = false;
    // a one-child pseudo group whose only child is the root
    Entry2l cur{root, (1u << (24 + oct)) | 1u};
    Entry2l leaf{0u, 0u};  // leaf set: record / triangle base, mask
    int sp = 0;
    for (;;) {
        // the leaf set of the last node: instances to enter, or triangles
        if (leaf.bits) {
            const int b = __builtin_ctz(leaf.bits);
            leaf.bits &= leaf.bits - 1u;
            if (top) {
                const TlasInstance& ti = insts[leaf.base + (uint32_t)b];
                // the top-level group, then a marker with the instances left
                // of this set (popping it returns to the world ray)
                if (cur.bits & 0xff000000u) stk.store(sp++, cur);
                stk.store(sp++, Entry2l{leaf.base, kMarker2l | leaf.bits});
                o = tlas_point(ti, wo);
                d = tlas_vector(ti, wd);
                inv = safe_inv_dir(d);
                sh = make_shear(d);
                oct = dir_octant(d);
                inst = ti.inst;
                top = false;
                cur = Entry2l{ti.blas_root, (1u << (24 + oct)) | 1u};
                leaf = Entry2l{0u, 0u};
                continue;
            }
            const BvhTri& tr = tris[leaf.base + (uint32_t)b];
            float t, u, v, det;
            if (intersect_triangle_uvd(o, sh, vec3(tr.v0x, tr.v0y, tr.v0z), vec3(tr.v1x, tr.v1y, tr.v1z),
                                       vec3(tr.v2x, tr.v2y, tr.v2z), tmin, tmax, t, u, v, det) &&
                hit_before(t, inst, tr.prim, hit.t, hit.inst, hit.prim)) {
                hit.t = t;
                hit.inst = inst;
                hit.prim = tr.prim;
                hit.u = u;
                hit.v = v;
                hit.det = det;
                hit.found = true;
                if (ANY) return true;
            }
            continue;
        }
        if (!(cur.bits & 0xff000000u)) {
            if (sp == 0) return hit.found;
            const Entry2l e = stk.load(--sp);
            if (e.bits & kMarker2l) {
                // back in the world: restore the ray, then the set's other instances
                o = wo;
                d = wd;
                inv = safe_inv_dir(d);
                sh = make_shear(d);
                oct = dir_octant(d);
                top = true;
                leaf = Entry2l{e.base, e.bits & ~kMarker2l};
                cur = Entry2l{0u, 0u};
                continue;
            }
            cur = e;
        }
        // next hit child of the group, in visit order
        const uint32_t hits = cur.bits >> 24;
        const int k = __builtin_ctz(hits);
        cur.bits &= ~(1u << (24 + k));
        const int slot = k ^ oct;
        const uint32_t idx = cur.base + (uint32_t)__builtin_popcount(cur.bits & 0xffu & ((1u << slot) - 1u));
        const Bvh8Node& n = nodes[idx];
        if (cur.bits & 0xff000000u) {
            // the group keeps its remaining hits on the stack
            stk.store(sp++, cur);
        }
        uint32_t ihits = 0, lbits = 0;
        const float far = hit.t * kBoxSlack;
        for (int c = 0; c < 8; c++) {
            const uint32_t meta = n.meta[c];
            if (meta == 0u) continue;
            vec3 lo, hi;
            child_box(n, c, lo, hi);
            float tn;
            if (!box_hit(lo, hi, o, inv, tmin, far, tn)) continue;
            if ((meta & 31u) >= 24u) {
                ihits |= 1u << c;  // internal child in slot c (offset 24 + c)
            } else {
                lbits |= (meta >> 5) << (meta & 31u);  // leaf span
            }
        }
        // internal hits from slot order (bit c) to visit order (bit c ^ oct)
        uint32_t v = 0;
        for (int c = 0; c < 8; c++)
            if (ihits & (1u << c)) v |= 1u << (c ^ oct);
        cur = Entry2l{n.child_base, (v << 24) | n.imask};
        leaf = Entry2l{n.tri_base, lbits};
    }
}

}  // namespace rm
