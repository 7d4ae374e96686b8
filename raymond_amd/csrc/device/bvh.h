// Compressed 8-wide BVH node and leaf-triangle records (this library's layout).
//
// Node: 80 bytes (five 16-byte loads).  Child boxes are quantised to 8 bits
// per axis on a per-node grid: decoded = p + q * 2^(e-127), with q*2^(e-127)
// exact and a single rounding in the addition.  The builder checks every
// decoded box against the exact child box and widens q until it contains it,
// so the quantised tree is conservative.  The design follows the compressed
// wide BVH of Ylitie, Karras and Laine (HPG 2017): 8-bit child boxes,
// per-child meta bytes, internal children stored contiguously and located by
// a popcount over `imask`.
//
// meta[i]:  0                      empty slot
//           0b001_xxxxx, xxxxx=24+i internal child in slot i
//           (2^n - 1) << 5 | off    leaf with n in 1..3 triangles at
//                                   tri_base + off (off + n <= 24)
// Child slots are assigned so that slot k holds the child nearest to corner k
// of the node (bit a of k set = high side on axis a); a ray with direction
// octant `oct` visits internal children in order slot = k ^ oct, k = 0..7.
#pragma once
#include "common.h"

namespace rm {

struct alignas(16) Bvh8Node {
    float px, py, pz;
    uint8_t ex, ey, ez, imask;
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t meta[8];
    uint8_t qlo_x[8], qlo_y[8], qlo_z[8];
    uint8_t qhi_x[8], qhi_y[8], qhi_z[8];
};
static_assert(sizeof(Bvh8Node) == 80, "Bvh8Node must be 80 bytes");

// Leaf triangle: world-space vertices, (instance, shape-local primitive) in w.
struct alignas(16) BvhTri {
    float v0x, v0y, v0z; uint32_t inst;
    float v1x, v1y, v1z; uint32_t prim;
    float v2x, v2y, v2z; uint32_t pad;
};
static_assert(sizeof(BvhTri) == 48, "BvhTri must be 48 bytes");

RM_HD float exp_scale(uint8_t e) { return u2f((uint32_t)e << 23); }
RM_HD float decode_q(float p, uint8_t q, float scale) { return p + (float)q * scale; }

// Position of point p on the 8 corners in octant order.
RM_HD int dir_octant(vec3 d) {
    return (d.x < 0.0f ? 1 : 0) | (d.y < 0.0f ? 2 : 0) | (d.z < 0.0f ? 4 : 0);
}

}  // namespace rm
