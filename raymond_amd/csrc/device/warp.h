// Sample warps and the orthonormal basis (device/utils/warp.hpp:1-59,
// device/utils/math.hpp:4-14).
#pragma once
#include "common.h"
#include "elem.h"

namespace rm {

// Frisvad-style basis, "Building an Orthonormal Basis, Revisited"
// (Duff et al. 2017); columns are (t, b, n) (device/utils/math.hpp:4-14).
RM_HD mat3 build_onb(vec3 n) {
    const float sign = copysign_(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    mat3 f;
    f.c0 = vec3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f.c1 = vec3(b, sign + n.y * n.y * a, -n.y);
    f.c2 = n;
    return f;
}

namespace warp {

// (device/utils/warp.hpp:5-10)
RM_HD vec2 equirect_sphere_to_square(vec3 v) {
    return vec2((atan2_(v.x, v.y) + kPi / 2.0f) / (2.0f * kPi), acos_(v.z / length(v)) / kPi);
}

// (device/utils/warp.hpp:12-20)
RM_HD vec3 uniform_square_to_sphere(vec2 uv) {
    const float z = 1.0f - 2.0f * uv.y;
    const float r = safe_sqrt(1.0f - z * z);
    float c;
    const float s = sincos_(2.0f * kPi * uv.x, c);
    return vec3(r * c, r * s, z);
}

// (device/utils/warp.hpp:22-26)
RM_HD vec2 uniform_sphere_to_square(vec3 v) {
    const float y = (1.0f - v.z) / 2.0f;
    const float x = atan2_(v.y, v.x) / (2.0f * kPi);
    return vec2(x < 0.0f ? x + 1.0f : x, y);
}

RM_HD float uniform_square_to_sphere_pdf() { return 1.0f / (4.0f * kPi); }

// (device/utils/warp.hpp:32-39)
RM_HD vec2 uniform_square_to_disk(vec2 uv) {
    float c;
    const float s = sincos_(2.0f * kPi * uv.x, c);
    const float r = sqrt_(uv.y);
    return vec2(r * c, r * s);
}

// cosine-weighted hemisphere, pdf = cos(theta)/pi (device/utils/warp.hpp:46-54)
RM_HD vec3 uniform_square_to_cosine_hemisphere(vec2 rnd) {
    const float cos_theta = sqrt_(rnd.x);
    const float sin_theta = sqrt_(1.0f - square(cos_theta));
    float cos_phi;
    const float sin_phi = sincos_(2.0f * kPi * rnd.y, cos_phi);
    return vec3(sin_theta * cos_phi, sin_theta * sin_phi, cos_theta);
}

// (device/utils/warp.hpp:56-59)
RM_HD vec2 uniform_square_to_triangle_barycentric(vec2 rnd) {
    const float x = sqrt_(rnd.x);
    return vec2(1.0f - x, x * rnd.y);
}

}  // namespace warp
}  // namespace rm
