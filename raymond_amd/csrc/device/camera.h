// Camera rays and the block-linear ray order of the reference's raygen
// (device/kernels/shading/generate.metal:72-121).
#pragma once
#include "common.h"
#include "rng.h"
#include "scene.h"
#include "spectral.h"
#include "elem.h"
#include "warp.h"

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


// ---------------------------------------------------------------------------
// Lens camera (generate.metal:16-70,122-163).  The reference traces through
// the un-vendored "lore" library; the surface model here is its own: a
// surface is a sphere cap of signed radius (0: a plane) with its vertex on
// the optical axis, `thickness` millimetres before the next surface toward
// the sensor, an aperture radius, and the medium that follows it toward the
// sensor given by n_d and the Abbe number V_d (V_d = 0: no dispersion).
// Surfaces are listed from the object side; the last one is the image plane
// and the first one only names the object-side medium.
struct LensSurface {
    float radius, thickness, aperture, ior_d, abbe;
};

struct Lens {
    const LensSurface* surfaces;
    int num_surfaces;        // 0: pinhole camera
    int stop_index;          // surface holding the aperture blades
    float relative_stop;     // opening of the stop, 0..1
    int aperture_blades;
    float sensor_scale;      // sensor size relative to 36 mm film
    float camera_scale;      // lens units (mm) to scene units
    float focus;             // sensor position along the axis (mm)
    int spectral;            // 1: one wavelength per path, weighted to RGB
};

// Two-term Cauchy dispersion n(l) = A + B / l^2 through n_d at the helium d
// line (587.56 nm) with the Abbe number V_d = (n_d - 1) / (n_F - n_C)
// (hydrogen F 486.13 nm, C 656.27 nm); l in micrometres.
RM_HD float lens_ior(const LensSurface& s, float l) {
    if (s.abbe <= 0.0f) return s.ior_d;
    const float lF = 0.48613f, lC = 0.65627f, ld = 0.58756f;
    const float B = (s.ior_d - 1.0f) / (s.abbe * (1.0f / (lF * lF) - 1.0f / (lC * lC)));
    return (s.ior_d - B / (ld * ld)) + B / (l * l);
}

// a mod 1 for a >= 0 (exact)
RM_HD float mod1(float a) { return a - (float)(int)a; }

// The stop's polygonal opening (generate.metal:24-34): `blades` straight
// edges that become a circle of radius relative_stop as the stop opens.
RM_HD bool lens_aperture(vec2 pos, const Lens& L) {
    const float phi = atan2_(pos.y, pos.x) / (2.0f * kPi) + (1.0f - L.relative_stop) / 4.0f;
    const float n = (float)L.aperture_blades;
    const float phi_l = mod1((phi + 1.0f) * n) - 0.5f;
    const float r = length(pos);
    const float x = kPi / n;
    const float norm = sqrt_(tan_(x) / x);
    const float blend = norm * cos_((2.0f * phi_l) * x);
    const float s2 = L.relative_stop * L.relative_stop;
    const float cool = r * ((1.0f - s2) * blend + s2);
    return cool <= L.relative_stop;
}

// Advances o along unit d to the surface whose vertex is at the origin;
// false if the ray misses the cap or passes outside the aperture.
RM_HD bool lens_propagate(vec3& o, vec3 d, const LensSurface& s, vec3& normal) {
    float t;
    if (s.radius == 0.0f) {
        if (d.z == 0.0f) return false;
        t = -o.z / d.z;
        normal = vec3(0.0f, 0.0f, 1.0f);
    } else {
        // centre on the axis at z = radius; the cap holds the vertex
        const vec3 c(0.0f, 0.0f, s.radius);
        const vec3 oc = o - c;
        const float b = dot(oc, d);
        const float disc = b * b - (dot(oc, oc) - s.radius * s.radius);
        if (disc < 0.0f) return false;
        const float q = sqrt_(disc);
        const float t0 = -b - q, t1 = -b + q;
        // of the two crossings, the one on the vertex's half of the sphere
        const float z0 = o.z + t0 * d.z;
        t = ((z0 - s.radius) * s.radius <= 0.0f) ? t0 : t1;
        normal = (o + t * d - c) * (1.0f / s.radius);
    }
    if (!(t >= 0.0f)) return false;  // a ray may start on the surface (the image plane)
    o = o + t * d;
    return o.x * o.x + o.y * o.y <= s.aperture * s.aperture;
}

// Snell refraction of unit d at unit normal n (either orientation) with
// eta = n_incident / n_transmitted; false on total internal reflection.
RM_HD bool lens_refract(vec3& d, vec3 n, float eta) {
    float c = -dot(d, n);
    if (c < 0.0f) {
        n = -n;
        c = -c;
    }
    const float k = 1.0f - eta * eta * (1.0f - c * c);
    if (k < 0.0f) return false;
    d = normalize(eta * d + (eta * c - sqrt_(k)) * n);
    return true;
}

// Traces a sensor ray through the surfaces toward the object side
// (generate.metal:37-69): at surface i the frame moves by its thickness to
// its vertex, the ray meets the cap, the stop tests the blades, and the ray
// refracts from the medium after surface i into the one before it.
RM_HD bool lens_trace(vec3& o, vec3& d, float lambda, const Lens& L) {
    int i = L.num_surfaces - 1;
    float n2 = lens_ior(L.surfaces[i], lambda);
    for (; i > 0; i--) {
        const LensSurface& s = L.surfaces[i];
        o.z += s.thickness;
        vec3 normal;
        if (!lens_propagate(o, d, s, normal)) return false;
        if (i == L.stop_index && !lens_aperture(vec2(o.x, o.y) * (1.0f / s.aperture), L)) return false;
        const float n1 = lens_ior(L.surfaces[i - 1], lambda);
        if (!lens_refract(d, normal, n2 / n1)) return false;
        n2 = n1;
    }
    return true;
}

// Camera ray through the lens for pixel (x, y) (generate.metal:106-157):
// two draws for the jitter, one for the wavelength, two for the point on
// the rear element.  The sensor point samples the rear element's disc by
// area; the weight is cos / pdf in solid angle, times the wavelength's RGB
// weight when spectral.  False when the ray does not leave the lens.
RM_HD bool lens_camera_ray(const Camera& cam, const Lens& L, uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                           Prng& prng, vec3& o, vec3& d, vec3& weight) {
    const vec2 r = prng.sample2d();
    const float u = ((((float)x + r.x) / (float)W + cam.shift.x) * 2.0f - 1.0f);
    const float v = -((((float)y + r.y) / (float)H + cam.shift.y) * 2.0f - 1.0f);
    const float aspect = (float)H / (float)W;
    const float t = prng.sample();
    const float lambda = (1.0f - t) * 0.38f + t * 0.78f;  // micrometres
    const LensSurface& rear = L.surfaces[L.num_surfaces - 2];
    const vec3 sensor(-u * L.sensor_scale * 36.0f / 2.0f, -v * L.sensor_scale * (36.0f * aspect) / 2.0f, L.focus);
    const vec2 disk = warp::uniform_square_to_disk(prng.sample2d());
    const vec3 aim(rear.aperture * disk.x, rear.aperture * disk.y, -rear.thickness);
    const vec3 du = aim - sensor;
    const float inv_d2 = 1.0f / length_squared(du);
    const vec3 dir = du * sqrt_(inv_d2);
    const float inv_pdf = (abs_(dir.z) * inv_d2) * (kPi * (rear.aperture * rear.aperture));
    o = sensor;
    d = dir;
    if (!lens_trace(o, d, lambda, L)) return false;
    const vec3 origin = L.camera_scale * o;
    o = cam.transform.point(origin);
    d = normalize(cam.transform.vector(d));
    const float w = abs_(dir.z) * inv_pdf;
    if (L.spectral) {
        const vec3 rgb = xyz_to_rgb(wavelength_to_xyz(1000.0f * lambda));
        weight = w * vec3(rgb.x * kCieNormRgbR, rgb.y * kCieNormRgbG, rgb.z * kCieNormRgbB) * 400.0f;
    } else {
        weight = vec3(w);
    }
    return true;
}

}  // namespace rm
