// Ray/triangle and ray/box tests shared by the GPU traversal kernels, the host
// BVH tools and the CPU oracle.
//
// The reference delegates intersection to Metal's closed intersector
// (device/kernels/shading/raytrace.metal:24-42), whose arithmetic is not
// published.  This library defines its own:
//   * triangles: the watertight test of Woop, Benthin and Wald (JCGT 2013):
//     shear to the dominant axis, edge functions in float with a double
//     re-evaluation when any is exactly zero, so no ray leaks between
//     triangles sharing an edge;
//   * boxes (the oracle's BVH2): slab test made conservative by scaling the
//     exit distance by 1 + 2*gamma(3) (Ize, "Robust BVH Ray Traversal", JCGT
//     2013), so a BVH can never cull a triangle that the exhaustive test
//     would hit.  The GPU traversal tests its quantised child boxes with its
//     own bounded-error form of the slab test and uses kBoxSlack on best_t.
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
    const float Az_ = comp(A, s.kz), Bz_ = comp(B, s.kz), Cz_ = comp(C, s.kz);
    const float Ax = comp(A, s.kx) - s.sx * Az_;
    const float Ay = comp(A, s.ky) - s.sy * Az_;
    const float Bx = comp(B, s.kx) - s.sx * Bz_;
    const float By = comp(B, s.ky) - s.sy * Bz_;
    const float Cx = comp(C, s.kx) - s.sx * Cz_;
    const float Cy = comp(C, s.ky) - s.sy * Cz_;
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
    const float Az = s.sz * Az_, Bz = s.sz * Bz_, Cz = s.sz * Cz_;
    const float T = (U * Az + V * Bz) + W * Cz;
    const float tt = T / det;
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

// 1 + 2*gamma(3) with gamma(n) = n u / (1 - n u), u = 2^-24, rounded up
constexpr float kBoxSlack = 1.00000048f;

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
