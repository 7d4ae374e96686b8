// Ray/triangle and ray/box tests shared by the GPU traversal kernels, the host
// BVH tools and the CPU oracle.
//
// The reference delegates intersection to Metal's closed intersector
// (device/kernels/shading/raytrace.metal:24-42), whose arithmetic is not
// published.  This library defines its own:
//   * triangles: the watertight test of Woop, Benthin and Wald (JCGT 2013):
//     shear to the dominant axis, edge functions in float with a double
//     re-evaluation when any is exactly zero, so no ray leaks between
//     triangles sharing an edge.  The distance is the plane distance,
//     within a few roundings of the exact one on planes normal to an axis
//     (ground planes, walls);
//   * boxes (the oracle's BVH2): slab test made conservative by scaling the
//     exit distance by 1 + 2^-20 (Ize, "Robust BVH Ray Traversal", JCGT 2013:
//     2*gamma(3) covers the slab arithmetic; the rest covers the triangle
//     distance's own rounding), so a BVH never culls a box holding a triangle
//     that the exhaustive test hits.  The GPU traversal tests its quantised
//     child boxes with its own bounded-error form of the slab test and uses
//     kBoxSlack on best_t.
// Closest-hit ties (equal t) go to the smaller (instance, primitive) pair, so
// the result does not depend on traversal order.
#pragma once
#include "common.h"

namespace rm {

struct RayShear {
    int kx, ky, kz;
    float sx, sy, sz;
};

RM_HD float comp(vec3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

RM_HD RayShear make_shear(vec3 d) {
    const float ax = abs_(d.x), ay = abs_(d.y), az = abs_(d.z);
    RayShear s;
    s.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    s.kx = s.kz == 2 ? 0 : s.kz + 1;
    s.ky = s.kx == 2 ? 0 : s.kx + 1;
    const float dz = comp(d, s.kz);
    if (dz < 0.0f) { const int t = s.kx; s.kx = s.ky; s.ky = t; }
    s.sx = comp(d, s.kx) / dz;
    s.sy = comp(d, s.ky) / dz;
    s.sz = 1.0f / dz;
    return s;
}

// Returns true when the triangle is hit at t with tmin < t < tmax, with the
// scaled barycentrics U, V and their sum det: v0 and v1 weigh U / det and
// V / det.  Traversal keeps these and divides once, for the closest hit.
RM_HD bool intersect_triangle_uvd(vec3 o, const RayShear& s, vec3 v0, vec3 v1, vec3 v2, float tmin, float tmax,
                                  float& t, float& u_out, float& v_out, float& det_out) {
    const vec3 A = v0 - o, B = v1 - o, C = v2 - o;
    // vertices relative to the origin in the permuted frame (kx, ky, kz)
    const vec3 a(comp(A, s.kx), comp(A, s.ky), comp(A, s.kz));
    const vec3 b(comp(B, s.kx), comp(B, s.ky), comp(B, s.kz));
    const vec3 c(comp(C, s.kx), comp(C, s.ky), comp(C, s.kz));
    const float Az_ = a.z, Bz_ = b.z, Cz_ = c.z;
    const float Ax = a.x - s.sx * Az_;
    const float Ay = a.y - s.sy * Az_;
    const float Bx = b.x - s.sx * Bz_;
    const float By = b.y - s.sy * Bz_;
    const float Cx = c.x - s.sx * Cz_;
    const float Cy = c.y - s.sy * Cz_;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        // products of floats are exact in double; one rounding in the difference
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = (U + V) + W;
    if (det == 0.0f) return false;
    // Distance to the triangle's plane, dot(n, a) / dot(n, d) with
    // n = (b - a) x (c - a) and d = (sx, sy, 1) / sz in this frame.  A
    // barycentric blend of the vertices' depths along the ray would carry
    // rounding of the order of u times the vertices' distance along the ray
    // (3e-6 for a 1e-4 hit on a 10-unit ground quad at a grazing angle), more
    // than the box tests' margins; on a plane normal to an axis n has two
    // exact zeros and this t is within a few roundings of the exact one.
    const vec3 n = cross(b - a, c - a);
    const float den = (n.x * s.sx + n.y * s.sy) + n.z;
    const float tt = (dot(n, a) * s.sz) / den;
    if (!(tt > tmin && tt < tmax)) return false;
    t = tt;
    u_out = U;
    v_out = V;
    det_out = det;
    return true;
}

// Returns true when the triangle is hit at t with tmin < t < tmax.
// b0, b1: barycentric weights of v0 and v1 (v2 gets 1 - b0 - b1).
RM_HD bool intersect_triangle(vec3 o, const RayShear& s, vec3 v0, vec3 v1, vec3 v2, float tmin, float tmax,
                              float& t, float& b0, float& b1) {
    float U, V, det;
    if (!intersect_triangle_uvd(o, s, v0, v1, v2, tmin, tmax, t, U, V, det)) return false;
    b0 = U / det;
    b1 = V / det;
    return true;
}

// closest-hit ordering: smaller t, then smaller instance, then smaller primitive
RM_HD bool hit_before(float t, uint32_t inst, uint32_t prim, float bt, uint32_t binst, uint32_t bprim) {
    if (t < bt) return true;
    if (t > bt) return false;
    if (inst != binst) return inst < binst;
    return prim < bprim;
}

// 1 + 16u with u = 2^-24: 2*gamma(3) (gamma(n) = n u / (1 - n u)) for the
// slab test's rounding and 6u for the triangle distance's, rounded up
constexpr float kBoxSlack = 1.00000095367431640625f;

// Reciprocal direction for the slab test; components below 1e-30 in
// magnitude are replaced by +-1e-30 so that 0 * inf never produces a NaN.
RM_HD vec3 safe_inv_dir(vec3 d) {
    auto f = [](float x) { return 1.0f / (abs_(x) > 1e-30f ? x : copysign_(1e-30f, x)); };
    return vec3(f(d.x), f(d.y), f(d.z));
}

// Conservative slab test against [lo, hi]; true when the box overlaps [tmin, tmax].
RM_HD bool box_hit(vec3 lo, vec3 hi, vec3 o, vec3 inv, float tmin, float tmax, float& tnear) {
    const float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
    const float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
    const float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
    const float tn = max_(max_(min_(tx0, tx1), min_(ty0, ty1)), max_(min_(tz0, tz1), tmin));
    const float tf = min_(min_(max_(tx0, tx1), max_(ty0, ty1)), min_(max_(tz0, tz1), tmax)) * kBoxSlack;
    tnear = tn;
    return tn <= tf;
}

}  // namespace rm
