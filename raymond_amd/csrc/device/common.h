// Shared host/device core: vector and matrix types, scalar helpers.
//
// Every header under device/ is compiled three ways from the same text:
//   * hipcc (ahead-of-time kernels for gfx950),
//   * hiprtc (the material/shading module, generated per scene),
//   * g++ (host code and the CPU oracle).
// All three builds use -ffp-contract=off and no fast-math, and every
// transcendental goes through elem.h, so a function evaluated on the host and
// on the GPU returns the same bits.  Arithmetic is written out explicitly
// (no implicit FMA, fixed association order) for that reason.
#pragma once

#if defined(__HIPCC_RTC__)
#define RM_HD __host__ __device__ inline __attribute__((always_inline))
#define RM_GPU 1
typedef unsigned char uint8_t;
typedef unsigned short uint16_t;
typedef unsigned int uint32_t;
typedef unsigned long long uint64_t;
typedef int int32_t;
typedef long long int64_t;
#elif defined(__HIPCC__)
#include <hip/hip_runtime.h>
#include <stdint.h>
#define RM_HD __host__ __device__ inline __attribute__((always_inline))
#define RM_GPU 1
#else
#include <stdint.h>
#include <string.h>
#include <math.h>
#define RM_HD inline
#endif

namespace rm {

// ---------------------------------------------------------------------------
// Bit casts and IEEE helpers (identical on host and device)

RM_HD uint32_t f2u(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
RM_HD float u2f(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }
RM_HD uint64_t d2u(double f) { uint64_t u; __builtin_memcpy(&u, &f, 8); return u; }
RM_HD double u2d(uint64_t u) { double f; __builtin_memcpy(&f, &u, 8); return f; }

RM_HD bool isnan_(float x) { return (f2u(x) & 0x7fffffffu) > 0x7f800000u; }
RM_HD bool isinf_(float x) { return (f2u(x) & 0x7fffffffu) == 0x7f800000u; }
RM_HD bool isfinite_(float x) { return (f2u(x) & 0x7f800000u) != 0x7f800000u; }
RM_HD float abs_(float x) { return u2f(f2u(x) & 0x7fffffffu); }
RM_HD float copysign_(float m, float s) { return u2f((f2u(m) & 0x7fffffffu) | (f2u(s) & 0x80000000u)); }
RM_HD double dabs_(double x) { return u2d(d2u(x) & 0x7fffffffffffffffull); }

// Explicit min/max with fixed NaN behaviour: returns b when the comparison fails.
RM_HD float min_(float a, float b) { return a < b ? a : b; }
RM_HD float max_(float a, float b) { return a > b ? a : b; }
RM_HD float saturate(float x) { return min_(max_(x, 0.0f), 1.0f); }
RM_HD float clamp_(float x, float lo, float hi) { return min_(max_(x, lo), hi); }
RM_HD int mini(int a, int b) { return a < b ? a : b; }
RM_HD int maxi(int a, int b) { return a > b ? a : b; }

// sign() as in Metal/GLSL: -1, 0 or +1 (0 for both zeros, NaN maps to 0).
RM_HD float sign_(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

RM_HD float sqrt_(float x) {
#if defined(RM_GPU)
    return __builtin_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
RM_HD double dsqrt_(double x) {
#if defined(RM_GPU)
    return __builtin_sqrt(x);
#else
    return sqrt(x);
#endif
}
RM_HD float floor_(float x) {
#if defined(RM_GPU)
    return __builtin_floorf(x);
#else
    return floorf(x);
#endif
}
RM_HD double dfloor_(double x) {
#if defined(RM_GPU)
    return __builtin_floor(x);
#else
    return floor(x);
#endif
}
RM_HD double drint_(double x) {
#if defined(RM_GPU)
    return __builtin_rint(x);
#else
    return rint(x);
#endif
}
RM_HD float fract_(float x) { return x - floor_(x); }

// safe_sqrt: 0 for non-positive input (device/utils/math.hpp:23)
RM_HD float safe_sqrt(float v) { return v <= 0.0f ? 0.0f : sqrt_(v); }
// safe_sqrtf: sqrt(max(f, 0)) (device/utils/math.hpp:16)
RM_HD float safe_sqrtf(float f) { return sqrt_(max_(f, 0.0f)); }
RM_HD float sqr(float f) { return f * f; }
RM_HD float square(float v) { return v * v; }
RM_HD float lerpf(float a, float b, float v) { return (1.0f - v) * a + v * b; }

// Ray flags (bridge/Ray.hpp:6-15)
enum RayFlags : uint32_t {
    RayCamera = 1u << 0,
    RayReflection = 1u << 1,
    RayTransmission = 1u << 2,
    RayShadow = 1u << 3,
    RayVolume = 1u << 4,
    RayDiffuse = 1u << 5,
    RayGlossy = 1u << 6,
    RaySingular = 1u << 7,
};

constexpr float kPi = 3.14159265358979323846f;      // M_PI_F
constexpr float kInvPi = 0.318309886183790671538f;  // M_1_PI_F
constexpr float kEps = 0.001f;                      // device/constants.hpp:3
constexpr float kInf = __builtin_huge_valf();

// ---------------------------------------------------------------------------
// Vectors

struct vec2 {
    float x, y;
    RM_HD vec2() : x(0), y(0) {}
    RM_HD vec2(float s) : x(s), y(s) {}
    RM_HD vec2(float a, float b) : x(a), y(b) {}
};
struct vec3 {
    float x, y, z;
    RM_HD vec3() : x(0), y(0), z(0) {}
    RM_HD vec3(float s) : x(s), y(s), z(s) {}
    RM_HD vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    RM_HD vec3(vec2 v, float c) : x(v.x), y(v.y), z(c) {}
    RM_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    RM_HD vec2 xy() const { return vec2(x, y); }
};
struct vec4 {
    float x, y, z, w;
    RM_HD vec4() : x(0), y(0), z(0), w(0) {}
    RM_HD vec4(float s) : x(s), y(s), z(s), w(s) {}
    RM_HD vec4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
    RM_HD vec4(vec3 v, float d) : x(v.x), y(v.y), z(v.z), w(d) {}
    RM_HD vec3 xyz() const { return vec3(x, y, z); }
};

// Component-wise selects on values (never on addresses).
RM_HD float select(bool c, float a, float b) { return c ? a : b; }
RM_HD vec3 select(bool c, vec3 a, vec3 b) {
    return vec3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z);
}
RM_HD vec4 select(bool c, vec4 a, vec4 b) {
    return vec4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w);
}

RM_HD vec2 operator+(vec2 a, vec2 b) { return vec2(a.x + b.x, a.y + b.y); }
RM_HD vec2 operator-(vec2 a, vec2 b) { return vec2(a.x - b.x, a.y - b.y); }
RM_HD vec2 operator*(vec2 a, vec2 b) { return vec2(a.x * b.x, a.y * b.y); }
RM_HD vec2 operator/(vec2 a, vec2 b) { return vec2(a.x / b.x, a.y / b.y); }
RM_HD vec2 operator*(vec2 a, float s) { return vec2(a.x * s, a.y * s); }
RM_HD vec2 operator*(float s, vec2 a) { return vec2(s * a.x, s * a.y); }
RM_HD vec2 operator/(vec2 a, float s) { return vec2(a.x / s, a.y / s); }
RM_HD vec2 operator+(vec2 a, float s) { return vec2(a.x + s, a.y + s); }
RM_HD vec2 operator-(vec2 a, float s) { return vec2(a.x - s, a.y - s); }
RM_HD vec2 operator-(vec2 a) { return vec2(-a.x, -a.y); }

RM_HD vec3 operator+(vec3 a, vec3 b) { return vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
RM_HD vec3 operator-(vec3 a, vec3 b) { return vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
RM_HD vec3 operator*(vec3 a, vec3 b) { return vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
RM_HD vec3 operator/(vec3 a, vec3 b) { return vec3(a.x / b.x, a.y / b.y, a.z / b.z); }
RM_HD vec3 operator*(vec3 a, float s) { return vec3(a.x * s, a.y * s, a.z * s); }
RM_HD vec3 operator*(float s, vec3 a) { return vec3(s * a.x, s * a.y, s * a.z); }
RM_HD vec3 operator/(vec3 a, float s) { return vec3(a.x / s, a.y / s, a.z / s); }
RM_HD vec3 operator/(float s, vec3 a) { return vec3(s / a.x, s / a.y, s / a.z); }
RM_HD vec3 operator+(vec3 a, float s) { return vec3(a.x + s, a.y + s, a.z + s); }
RM_HD vec3 operator+(float s, vec3 a) { return vec3(s + a.x, s + a.y, s + a.z); }
RM_HD vec3 operator-(vec3 a, float s) { return vec3(a.x - s, a.y - s, a.z - s); }
RM_HD vec3 operator-(float s, vec3 a) { return vec3(s - a.x, s - a.y, s - a.z); }
RM_HD vec3 operator-(vec3 a) { return vec3(-a.x, -a.y, -a.z); }
RM_HD vec3& operator+=(vec3& a, vec3 b) { a = a + b; return a; }
RM_HD vec3& operator-=(vec3& a, vec3 b) { a = a - b; return a; }
RM_HD vec3& operator*=(vec3& a, vec3 b) { a = a * b; return a; }
RM_HD vec3& operator*=(vec3& a, float s) { a = a * s; return a; }
RM_HD vec3& operator/=(vec3& a, float s) { a = a / s; return a; }

RM_HD vec4 operator+(vec4 a, vec4 b) { return vec4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
RM_HD vec4 operator-(vec4 a, vec4 b) { return vec4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
RM_HD vec4 operator*(vec4 a, vec4 b) { return vec4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
RM_HD vec4 operator*(vec4 a, float s) { return vec4(a.x * s, a.y * s, a.z * s, a.w * s); }
RM_HD vec4 operator*(float s, vec4 a) { return vec4(s * a.x, s * a.y, s * a.z, s * a.w); }
RM_HD vec4 operator+(vec4 a, float s) { return vec4(a.x + s, a.y + s, a.z + s, a.w + s); }
RM_HD vec4 operator-(vec4 a, float s) { return vec4(a.x - s, a.y - s, a.z - s, a.w - s); }
RM_HD vec4 operator-(float s, vec4 a) { return vec4(s - a.x, s - a.y, s - a.z, s - a.w); }

RM_HD float dot(vec2 a, vec2 b) { return a.x * b.x + a.y * b.y; }
RM_HD float dot(vec3 a, vec3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
RM_HD vec3 cross(vec3 a, vec3 b) {
    return vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
RM_HD float length_squared(vec3 v) { return dot(v, v); }
RM_HD float length(vec3 v) { return sqrt_(dot(v, v)); }
RM_HD float length(vec2 v) { return sqrt_(dot(v, v)); }
// normalize: v scaled by the correctly rounded reciprocal of its length.
RM_HD vec3 normalize(vec3 v) { return v * (1.0f / sqrt_(dot(v, v))); }
RM_HD vec3 vmin(vec3 a, vec3 b) { return vec3(min_(a.x, b.x), min_(a.y, b.y), min_(a.z, b.z)); }
RM_HD vec3 vmax(vec3 a, vec3 b) { return vec3(max_(a.x, b.x), max_(a.y, b.y), max_(a.z, b.z)); }
RM_HD vec3 saturate(vec3 v) { return vec3(saturate(v.x), saturate(v.y), saturate(v.z)); }
RM_HD vec4 saturate(vec4 v) { return vec4(saturate(v.x), saturate(v.y), saturate(v.z), saturate(v.w)); }
RM_HD vec3 vabs(vec3 v) { return vec3(abs_(v.x), abs_(v.y), abs_(v.z)); }
// mean of a float3 (device/utils/math.hpp:41)
RM_HD float mean(vec3 v) { return (v.x + v.y + v.z) / 3.0f; }
RM_HD bool all_zero(vec3 v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f; }
RM_HD bool any_nonzero(vec3 v) { return v.x != 0.0f || v.y != 0.0f || v.z != 0.0f; }
RM_HD bool all_finite(vec3 v) { return isfinite_(v.x) && isfinite_(v.y) && isfinite_(v.z); }
RM_HD vec3 lerp(vec3 a, vec3 b, float v) { return (1.0f - v) * a + v * b; }
RM_HD vec4 lerp(vec4 a, vec4 b, float v) { return (1.0f - v) * a + v * b; }
// reflect/refract with Metal's definitions (incident i points towards the surface).
RM_HD vec3 reflect(vec3 i, vec3 n) { return i - 2.0f * dot(n, i) * n; }
RM_HD vec3 refract(vec3 i, vec3 n, float eta) {
    const float ndi = dot(n, i);
    const float k = 1.0f - eta * eta * (1.0f - ndi * ndi);
    if (k < 0.0f) return vec3(0.0f);
    return eta * i - (eta * ndi + sqrt_(k)) * n;
}
// safe_divide with per-component fallback (device/utils/math.hpp:130-136)
RM_HD float safe_divide(float a, float b, float fallback) { return b == 0.0f ? fallback : a / b; }
RM_HD vec3 safe_divide(vec3 a, vec3 b, float fallback) {
    return vec3(safe_divide(a.x, b.x, fallback), safe_divide(a.y, b.y, fallback), safe_divide(a.z, b.z, fallback));
}

// ---------------------------------------------------------------------------
// Matrices, stored as columns (Metal's convention).

struct mat3 {
    vec3 c0, c1, c2;
    // M * v
    RM_HD vec3 operator*(vec3 v) const { return (c0 * v.x + c1 * v.y) + c2 * v.z; }
};
// v * M  (row vector times matrix) = (dot(v,c0), dot(v,c1), dot(v,c2))
RM_HD vec3 operator*(vec3 v, const mat3& m) { return vec3(dot(v, m.c0), dot(v, m.c1), dot(v, m.c2)); }

struct mat4 {
    vec4 c0, c1, c2, c3;
    // (M * (p, 1)).xyz
    RM_HD vec3 point(vec3 p) const {
        return vec3(((c0.x * p.x + c1.x * p.y) + c2.x * p.z) + c3.x,
                    ((c0.y * p.x + c1.y * p.y) + c2.y * p.z) + c3.y,
                    ((c0.z * p.x + c1.z * p.y) + c2.z * p.z) + c3.z);
    }
    // (M * (v, 0)).xyz
    RM_HD vec3 vector(vec3 v) const {
        return vec3((c0.x * v.x + c1.x * v.y) + c2.x * v.z,
                    (c0.y * v.x + c1.y * v.y) + c2.y * v.z,
                    (c0.z * v.x + c1.z * v.y) + c2.z * v.z);
    }
};

}  // namespace rm
