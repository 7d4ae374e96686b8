// Hashes and gradient noise used by the texture nodes (device/utils/noise.hpp:1-441).
// Hash: Bob Jenkins' lookup3 final mix (public domain, burtleburtle.net/bob/c/lookup3.c).
// Noise: improved Perlin gradient noise and fBm with the Cycles conventions
// (gradient tables, output scales, octave blending).
#pragma once
#include "common.h"

namespace rm {

RM_HD uint32_t rot32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

RM_HD void lookup3_mix(uint32_t& a, uint32_t& b, uint32_t& c) {
    a -= c; a ^= rot32(c, 4);  c += b;
    b -= a; b ^= rot32(a, 6);  a += c;
    c -= b; c ^= rot32(b, 8);  b += a;
    a -= c; a ^= rot32(c, 16); c += b;
    b -= a; b ^= rot32(a, 19); a += c;
    c -= b; c ^= rot32(b, 4);  b += a;
}
RM_HD void lookup3_final(uint32_t& a, uint32_t& b, uint32_t& c) {
    c ^= b; c -= rot32(b, 14);
    a ^= c; a -= rot32(c, 11);
    b ^= a; b -= rot32(a, 25);
    c ^= b; c -= rot32(b, 16);
    a ^= c; a -= rot32(c, 4);
    b ^= a; b -= rot32(a, 14);
    c ^= b; c -= rot32(b, 24);
}

RM_HD uint32_t hash_uint(uint32_t kx) {
    uint32_t a, b, c;
    a = b = c = 0xdeadbeefu + (1u << 2) + 13u;
    a += kx;
    lookup3_final(a, b, c);
    return c;
}
RM_HD uint32_t hash_uint2(uint32_t kx, uint32_t ky) {
    uint32_t a, b, c;
    a = b = c = 0xdeadbeefu + (2u << 2) + 13u;
    b += ky;
    a += kx;
    lookup3_final(a, b, c);
    return c;
}
RM_HD uint32_t hash_uint3(uint32_t kx, uint32_t ky, uint32_t kz) {
    uint32_t a, b, c;
    a = b = c = 0xdeadbeefu + (3u << 2) + 13u;
    c += kz;
    b += ky;
    a += kx;
    lookup3_final(a, b, c);
    return c;
}
RM_HD uint32_t hash_uint4(uint32_t kx, uint32_t ky, uint32_t kz, uint32_t kw) {
    uint32_t a, b, c;
    a = b = c = 0xdeadbeefu + (4u << 2) + 13u;
    a += kx;
    b += ky;
    c += kz;
    lookup3_mix(a, b, c);
    a += kw;
    lookup3_final(a, b, c);
    return c;
}

RM_HD float hash_uint_to_float(uint32_t kx) { return (float)hash_uint(kx) / (float)0xFFFFFFFFu; }
RM_HD float hash_uint2_to_float(uint32_t kx, uint32_t ky) { return (float)hash_uint2(kx, ky) / (float)0xFFFFFFFFu; }
RM_HD float hash_float_to_float(float k) { return hash_uint_to_float(f2u(k)); }
RM_HD float hash_float2_to_float(vec2 k) { return hash_uint2_to_float(f2u(k.x), f2u(k.y)); }

// (int)x minus one for negative x -- note: -1.0 maps to -2 as in the reference
RM_HD int quick_floor_to_int(float x) { return (int)x - ((x < 0.0f) ? 1 : 0); }
RM_HD float floorfrac(float x, int& i) {
    i = quick_floor_to_int(x);
    return x - (float)i;
}

RM_HD float fade(float t) { return ((t * t) * t) * (t * (t * 6.0f - 15.0f) + 10.0f); }
// v with its sign flipped when bit b of h is set: the bits of
// ((h >> b) & 1) ? -v : v, by an XOR into the sign bit instead of a compare
// and a select (on gfx950 the select waits on the compare's VCC write).
#ifndef RM_NOISE_XOR_SIGN
#define RM_NOISE_XOR_SIGN 1
#endif
RM_HD float negate_bit(float v, uint32_t h, int b) {
#if RM_NOISE_XOR_SIGN
    return u2f(f2u(v) ^ ((h << (31 - b)) & 0x80000000u));
#else
    return ((h >> b) & 1u) ? -v : v;
#endif
}
RM_HD float mixf(float a, float b, float t) { return a + (b - a) * t; }

RM_HD float grad1(int hash, float x) {
    const int h = hash & 15;
    const float g = (float)(1 + (h & 7));
    return negate_bit(g, (uint32_t)h, 3) * x;
}
RM_HD float grad2(int hash, float x, float y) {
    const int h = hash & 7;
    const float u = h < 4 ? x : y;
    const float v = 2.0f * (h < 4 ? y : x);
    return negate_bit(u, (uint32_t)h, 0) + negate_bit(v, (uint32_t)h, 1);
}
RM_HD float grad3(int hash, float x, float y, float z) {
    const int h = hash & 15;
    const float u = h < 8 ? x : y;
    const float vt = ((h == 12) || (h == 14)) ? x : z;
    const float v = h < 4 ? y : vt;
    return negate_bit(u, (uint32_t)h, 0) + negate_bit(v, (uint32_t)h, 1);
}
RM_HD float grad4(int hash, float x, float y, float z, float w) {
    const int h = hash & 31;
    const float u = h < 24 ? x : y;
    const float v = h < 16 ? y : z;
    const float s = h < 8 ? z : w;
    return (negate_bit(u, (uint32_t)h, 0) + negate_bit(v, (uint32_t)h, 1)) + negate_bit(s, (uint32_t)h, 2);
}

RM_HD float bi_mix(float v0, float v1, float v2, float v3, float x, float y) {
    const float x1 = 1.0f - x;
    return (1.0f - y) * (v0 * x1 + v1 * x) + y * (v2 * x1 + v3 * x);
}
RM_HD float tri_mix(float v0, float v1, float v2, float v3, float v4, float v5, float v6, float v7, float x,
                    float y, float z) {
    const float x1 = 1.0f - x, y1 = 1.0f - y, z1 = 1.0f - z;
    return z1 * (y1 * (v0 * x1 + v1 * x) + y * (v2 * x1 + v3 * x)) +
           z * (y1 * (v4 * x1 + v5 * x) + y * (v6 * x1 + v7 * x));
}

RM_HD float perlin_1d(float x) {
    int X;
    const float fx = floorfrac(x, X);
    const float u = fade(fx);
    return mixf(grad1((int)hash_uint((uint32_t)X), fx), grad1((int)hash_uint((uint32_t)(X + 1)), fx - 1.0f), u);
}
RM_HD float perlin_2d(float x, float y) {
    int X, Y;
    const float fx = floorfrac(x, X), fy = floorfrac(y, Y);
    const float u = fade(fx), v = fade(fy);
    auto H = [](int a, int b) { return (int)hash_uint2((uint32_t)a, (uint32_t)b); };
    return bi_mix(grad2(H(X, Y), fx, fy), grad2(H(X + 1, Y), fx - 1.0f, fy), grad2(H(X, Y + 1), fx, fy - 1.0f),
                  grad2(H(X + 1, Y + 1), fx - 1.0f, fy - 1.0f), u, v);
}
RM_HD float perlin_3d(float x, float y, float z) {
    int X, Y, Z;
    const float fx = floorfrac(x, X), fy = floorfrac(y, Y), fz = floorfrac(z, Z);
    const float u = fade(fx), v = fade(fy), w = fade(fz);
    auto H = [](int a, int b, int c) { return (int)hash_uint3((uint32_t)a, (uint32_t)b, (uint32_t)c); };
    return tri_mix(grad3(H(X, Y, Z), fx, fy, fz), grad3(H(X + 1, Y, Z), fx - 1.0f, fy, fz),
                   grad3(H(X, Y + 1, Z), fx, fy - 1.0f, fz), grad3(H(X + 1, Y + 1, Z), fx - 1.0f, fy - 1.0f, fz),
                   grad3(H(X, Y, Z + 1), fx, fy, fz - 1.0f), grad3(H(X + 1, Y, Z + 1), fx - 1.0f, fy, fz - 1.0f),
                   grad3(H(X, Y + 1, Z + 1), fx, fy - 1.0f, fz - 1.0f),
                   grad3(H(X + 1, Y + 1, Z + 1), fx - 1.0f, fy - 1.0f, fz - 1.0f), u, v, w);
}

// four-linear blend of two tri_mix cubes along w (device/utils/noise.hpp:218-241)
RM_HD float quad_mix(float v0, float v1, float v2, float v3, float v4, float v5, float v6, float v7, float v8,
                     float v9, float v10, float v11, float v12, float v13, float v14, float v15, float x, float y,
                     float z, float w) {
    return mixf(tri_mix(v0, v1, v2, v3, v4, v5, v6, v7, x, y, z), tri_mix(v8, v9, v10, v11, v12, v13, v14, v15, x, y, z),
                w);
}
// (device/utils/noise.hpp:313-353): the 16 corners of the unit tesseract,
// x fastest, then y, z and w
RM_HD float perlin_4d(float x, float y, float z, float w) {
    int X, Y, Z, W;
    const float fx = floorfrac(x, X), fy = floorfrac(y, Y), fz = floorfrac(z, Z), fw = floorfrac(w, W);
    const float u = fade(fx), v = fade(fy), t = fade(fz), s = fade(fw);
    auto G = [&](int dx, int dy, int dz, int dw) {
        return grad4((int)hash_uint4((uint32_t)(X + dx), (uint32_t)(Y + dy), (uint32_t)(Z + dz), (uint32_t)(W + dw)),
                     fx - (float)dx, fy - (float)dy, fz - (float)dz, fw - (float)dw);
    };
    return quad_mix(G(0, 0, 0, 0), G(1, 0, 0, 0), G(0, 1, 0, 0), G(1, 1, 0, 0), G(0, 0, 1, 0), G(1, 0, 1, 0),
                    G(0, 1, 1, 0), G(1, 1, 1, 0), G(0, 0, 0, 1), G(1, 0, 0, 1), G(0, 1, 0, 1), G(1, 1, 0, 1),
                    G(0, 0, 1, 1), G(1, 0, 1, 1), G(0, 1, 1, 1), G(1, 1, 1, 1), u, v, t, s);
}

// IEEE-only finiteness test (device/utils/noise.hpp:387-392)
RM_HD bool isfinite_safe(float f) {
    const uint32_t x = f2u(f);
    return (f == f) && (x == 0 || x == (1u << 31) || (f != 2.0f * f)) && !((x << 1) > 0xff000000u);
}
RM_HD float ensure_finite(float v) { return isfinite_safe(v) ? v : 0.0f; }

RM_HD float snoise1(float p) { return 0.2500f * ensure_finite(perlin_1d(p)); }
RM_HD float snoise2(vec2 p) { return 0.6616f * ensure_finite(perlin_2d(p.x, p.y)); }
RM_HD float snoise3(vec3 p) { return 0.9820f * ensure_finite(perlin_3d(p.x, p.y, p.z)); }
RM_HD float snoise4(vec4 p) { return 0.8344f * ensure_finite(perlin_4d(p.x, p.y, p.z, p.w)); }
RM_HD float noise1(float p) { return 0.5f * snoise1(p) + 0.5f; }
RM_HD float noise2(vec2 p) { return 0.5f * snoise2(p) + 0.5f; }
RM_HD float noise3(vec3 p) { return 0.5f * snoise3(p) + 0.5f; }
RM_HD float noise4(vec4 p) { return 0.5f * snoise4(p) + 0.5f; }

RM_HD float noise_any(float p) { return noise1(p); }
RM_HD float noise_any(vec2 p) { return noise2(p); }
RM_HD float noise_any(vec3 p) { return noise3(p); }
RM_HD float noise_any(vec4 p) { return noise4(p); }

// fBm with fractional octave blending (device/utils/noise.hpp:419-441)
template <typename T>
RM_HD float fractal_noise(T p, float octaves, float roughness) {
    float fscale = 1.0f, amp = 1.0f, maxamp = 0.0f, sum = 0.0f;
    octaves = clamp_(octaves, 0.0f, 15.0f);
    const int n = (int)octaves;
    for (int i = 0; i <= n; i++) {
        const float t = noise_any(fscale * p);
        sum += t * amp;
        maxamp += amp;
        amp *= clamp_(roughness, 0.0f, 1.0f);
        fscale *= 2.0f;
    }
    const float rmd = octaves - floor_(octaves);
    if (rmd != 0.0f) {
        const float t = noise_any(fscale * p);
        float sum2 = sum + t * amp;
        sum /= maxamp;
        sum2 /= maxamp + amp;
        return (1.0f - rmd) * sum + rmd * sum2;
    }
    return sum / maxamp;
}

}  // namespace rm
