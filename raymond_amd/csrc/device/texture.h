// Texture lookups of the material library: bilinear filtering with repeat
// addressing over normalised coordinates, the sampler the reference's image
// and sky nodes declare (device/nodes/nodes.hpp:517-518,831-832;
// Metal's `filter::linear, address::repeat, coord::normalized`).  Texel
// (i, j) covers [i, i+1) x [j, j+1) / (width, height) and its value sits at
// its centre; row 0 is v in [0, 1/height).
#pragma once
#include "common.h"
#include "scene.h"

namespace rm {

RM_HD int wrap_index(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

RM_HD vec4 texel(const Texture& t, int x, int y) {
    const float* p = t.pixels + (size_t)t.channels * ((size_t)y * (size_t)t.width + (size_t)x);
    // a one-channel image reads as grey with alpha 1 (the reference's
    // PIXEL_FORMAT_R, nodes.hpp:528-535)
    return t.channels == 1 ? vec4(p[0], p[0], p[0], 1.0f) : vec4(p[0], p[1], p[2], p[3]);
}

RM_HD vec4 texture_sample(const Texture& t, vec2 uv) {
    const float fx = uv.x * (float)t.width - 0.5f;
    const float fy = uv.y * (float)t.height - 0.5f;
    const float x0f = floor_(fx), y0f = floor_(fy);
    const float ax = fx - x0f, ay = fy - y0f;
    // coordinates far outside [0, 1] wrap through the float before the int
    const int x0 = wrap_index((int)(x0f - floor_(x0f / (float)t.width) * (float)t.width), t.width);
    const int y0 = wrap_index((int)(y0f - floor_(y0f / (float)t.height) * (float)t.height), t.height);
    const int x1 = wrap_index(x0 + 1, t.width), y1 = wrap_index(y0 + 1, t.height);
    const vec4 a = texel(t, x0, y0), b = texel(t, x1, y0), c = texel(t, x0, y1), d = texel(t, x1, y1);
    return (a * (1.0f - ax) + b * ax) * (1.0f - ay) + (c * (1.0f - ax) + d * ax) * ay;
}

}  // namespace rm
