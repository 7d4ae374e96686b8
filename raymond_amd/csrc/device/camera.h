// Camera rays and the block-linear ray order of the reference's raygen
// (device/kernels/shading/generate.metal:72-121).
#pragma once
#include "common.h"
#include "rng.h"
#include "scene.h"

namespace rm {

// Inverse of the reference's block-linear index over 8x8 threadgroups
// (generate.metal:89-91):
//   rayIndex = tx + ty*aw.x + gx*8*aw.y + gy*8*W,
// where aw = actual threadgroup size (smaller at the right/bottom edge).
RM_HD void block_linear_to_pixel(uint32_t ray_index, uint32_t W, uint32_t H, uint32_t& x, uint32_t& y) {
    const uint32_t row_span = 8u * W;
    const uint32_t gy = ray_index / row_span;
    const uint32_t h_g = (H - 8u * gy) < 8u ? (H - 8u * gy) : 8u;
    const uint32_t off = ray_index - gy * row_span;
    const uint32_t gx = off / (8u * h_g);
    const uint32_t w_g = (W - 8u * gx) < 8u ? (W - 8u * gx) : 8u;
    const uint32_t r = off - gx * 8u * h_g;
    x = 8u * gx + r % w_g;
    y = 8u * gy + r / w_g;
}

// Pinhole camera ray for pixel (x, y); consumes two RNG draws for the
// sub-pixel jitter (generate.metal:108-121).
RM_HD void camera_ray(const Camera& cam, uint32_t x, uint32_t y, uint32_t W, uint32_t H, Prng& prng, vec3& o,
                      vec3& d) {
    const vec2 r = prng.sample2d();
    const float jx = (float)x + r.x;
    const float jy = (float)y + r.y;
    const float u = ((jx / (float)W + cam.shift.x) * 2.0f - 1.0f);
    const float v = -((jy / (float)H + cam.shift.y) * 2.0f - 1.0f);
    const float aspect = (float)H / (float)W;
    o = cam.transform.point(vec3(0.0f));
    d = normalize(cam.transform.vector(vec3(u, v * aspect, -cam.focal_length)));
}

}  // namespace rm
