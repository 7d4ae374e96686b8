// Nishita sky model: the precomputed sky texture and sun colours that the
// reference's TexNishita node samples (host/sky/sky_nishita.cpp via
// SkyLoader.mm:8-39, used by MaterialBuilder.swift:329-378).
//
// Single scattering in a spherical atmosphere (Nishita, Sirai, Tadamura and
// Nakamae, SIGGRAPH 1993) with the parameters of Blender's "Nishita" sky:
// Rayleigh and Mie layers with exponential density, an ozone layer with a
// tent profile between 10 and 40 km, 21 wavelengths from 380 to 780 nm
// integrated against the CIE 1931 colour matching functions.  Along the view
// ray 32 segments are sampled at their midpoints; the optical depth toward
// the sun uses 8-point Gauss-Laguerre quadrature over the light ray scaled so
// that its last node lies inside the atmosphere (the exponential weight is
// folded into the weights).  Results are CIE XYZ.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../../../include/rm_host.h"

namespace rm {
namespace sky {
namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kEarthRadius = 6360e3f;       // m
constexpr float kAtmosphereRadius = 6420e3f;  // m
constexpr float kRayleighHeight = 8e3f;       // scale height (m)
constexpr float kMieHeight = 1.2e3f;          // scale height (m)
constexpr float kMieScattering = 2e-5f;       // m^-1
constexpr float kMieExtinction = 1.11f;       // extinction / scattering of the aerosols
constexpr float kMieG = 0.76f;                // aerosol anisotropy
constexpr int kSegments = 32;
constexpr int kBands = 21;
constexpr float kBandWidth = 20.0f;           // nm between sampled wavelengths

// Per 20-nm band from 380 nm: solar irradiance at the top of the atmosphere
// (W m^-2 nm^-1), Rayleigh scattering and ozone absorption coefficients
// (m^-1) and the CIE 1931 2-degree colour matching functions x, y, z.
struct Band {
    float irradiance, rayleigh, ozone, x, y, z;
};
const Band kBandTable[kBands] = {
    {1.45756829855592995315f, 0.00005424820087636473f, 0.00000000325126849861f, 0.00136800000f, 0.00003900000f, 0.00645000100f},
    {1.56596305559738380175f, 0.00004418549866505454f, 0.00000000585395365047f, 0.01431000000f, 0.00039600000f, 0.06785001000f},
    {1.65148449067670455293f, 0.00003635151910165377f, 0.00000001977191155085f, 0.13438000000f, 0.00400000000f, 0.64560000000f},
    {1.71496242737209314555f, 0.00003017929012024763f, 0.00000007309568762914f, 0.34828000000f, 0.02300000000f, 1.74706000000f},
    {1.75797983805020541226f, 0.00002526320226989157f, 0.00000020084561514287f, 0.29080000000f, 0.06000000000f, 1.66920000000f},
    {1.78256407885924539336f, 0.00002130859310621843f, 0.00000040383958096161f, 0.09564000000f, 0.13902000000f, 0.81295010000f},
    {1.79095108475838560302f, 0.00001809838025320633f, 0.00000063551335912363f, 0.00490000000f, 0.32300000000f, 0.27200000000f},
    {1.78541550133410664714f, 0.00001547057129129042f, 0.00000096707041180970f, 0.06327000000f, 0.71000000000f, 0.07824999000f},
    {1.76815554864306845317f, 0.00001330284977336850f, 0.00000154797400424410f, 0.29040000000f, 0.95400000000f, 0.02030000000f},
    {1.74122069647250410362f, 0.00001150184784075764f, 0.00000209038647223331f, 0.59450000000f, 0.99500000000f, 0.00390000000f},
    {1.70647127164943679389f, 0.00000999557429990163f, 0.00000246128056164565f, 0.91630000000f, 0.87000000000f, 0.00165000100f},
    {1.66556087452739887134f, 0.00000872799973630707f, 0.00000273551299461512f, 1.06220000000f, 0.63100000000f, 0.00080000000f},
    {1.61993437242451854274f, 0.00000765513700977967f, 0.00000215125863128643f, 0.85444990000f, 0.38100000000f, 0.00019000000f},
    {1.57083597368892080581f, 0.00000674217203751443f, 0.00000159051840791988f, 0.44790000000f, 0.17500000000f, 0.00002000000f},
    {1.51932335059305478886f, 0.00000596134125832052f, 0.00000112356197979857f, 0.16490000000f, 0.06100000000f, 0.00000000000f},
    {1.46628494965214395407f, 0.00000529034598065810f, 0.00000073527551487574f, 0.04677000000f, 0.01700000000f, 0.00000000000f},
    {1.41245852740172450623f, 0.00000471115687557433f, 0.00000046450130357806f, 0.01135916000f, 0.00410200000f, 0.00000000000f},
    {1.35844961970384092709f, 0.00000420910481110487f, 0.00000033096079921048f, 0.00289932700f, 0.00104700000f, 0.00000000000f},
    {1.30474913844739281998f, 0.00000377218381260133f, 0.00000022512612292678f, 0.00069007860f, 0.00024920000f, 0.00000000000f},
    {1.25174963272610817455f, 0.00000339051255477280f, 0.00000014879129266490f, 0.00016615050f, 0.00006000000f, 0.00000000000f},
    {1.19975998755420620867f, 0.00000305591531679811f, 0.00000016828623364192f, 0.00004150994f, 0.00001499000f, 0.00000000000f},
};

struct V3 {
    float x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float norm(V3 a) { return std::sqrt(dot(a, a)); }

// 8-point Gauss-Laguerre rule for the light ray: nodes x_i / 25 and weights
// w_i exp(x_i) / 25, so that sum_i w'_i f(x'_i L) L approximates the
// integral of f over [0, L] for f that decays roughly exponentially.  The
// nodes are the roots of the Laguerre polynomial L_8, found by Newton's
// method in double precision.
struct Quadrature {
    float node[8], weight[8];
    Quadrature() {
        const int n = 8;
        double x = 0.0;
        for (int i = 0; i < n; i++) {
            // starting guesses of Stroud and Secrest for the Laguerre roots
            if (i == 0) x = 3.0 / (1.0 + 2.4 * n);
            else if (i == 1) x += 15.0 / (1.0 + 2.5 * n);
            else x += (1.0 + 2.55 * (i - 1)) / (1.9 * (i - 1)) * (x - prev_[i - 2]);
            double p1 = 0.0, p2 = 0.0;
            for (int it = 0; it < 100; it++) {
                p1 = 1.0;
                p2 = 0.0;
                for (int j = 0; j < n; j++) {  // three-term recurrence for L_n(x)
                    const double p3 = p2;
                    p2 = p1;
                    p1 = ((2.0 * j + 1.0 - x) * p2 - j * p3) / (j + 1.0);
                }
                const double dp = n * (p1 - p2) / x;
                const double dx = p1 / dp;
                x -= dx;
                if (std::fabs(dx) < 1e-15 * x) break;
            }
            prev_[i] = x;
            // w = x / ((n + 1)^2 L_{n+1}(x)^2) = 1 / (x L_n'(x)^2)
            double l = 1.0, lm = 0.0;
            for (int j = 0; j < n; j++) {
                const double t = lm;
                lm = l;
                l = ((2.0 * j + 1.0 - x) * lm - j * t) / (j + 1.0);
            }
            const double dl = n * (l - lm) / x;
            const double w = 1.0 / (x * dl * dl);
            node[i] = (float)(x / 25.0);
            weight[i] = (float)(w * std::exp(x) / 25.0);
        }
    }

private:
    double prev_[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

const Quadrature& quadrature() {
    static const Quadrature q;
    return q;
}

// density factors (Rayleigh, Mie, ozone) at `height` metres above the ground
V3 density(float height) {
    float ozone = 0.0f;
    if (height >= 10000.0f && height < 25000.0f) ozone = height / 15000.0f - 2.0f / 3.0f;
    else if (height >= 25000.0f && height < 40000.0f) ozone = 8.0f / 3.0f - height / 15000.0f;
    return {std::exp(-height / kRayleighHeight), std::exp(-height / kMieHeight), ozone};
}

// distance from p (inside the atmosphere) along unit d to the atmosphere's shell
float to_space(V3 p, V3 d) {
    const float b = 2.0f * dot(d, p);
    const float c = dot(p, p) - kAtmosphereRadius * kAtmosphereRadius;
    return (-b + std::sqrt(b * b - 4.0f * c)) / 2.0f;
}

// length of the segment from p along unit d to the atmosphere's shell
float path_length(V3 p, V3 d) { return norm(p + to_space(p, d) * d - p); }

// whether the ray from p along d meets the ground
bool hits_ground(V3 p, V3 d) {
    if (d.z >= 0.0f) return false;
    const float b = 2.0f * dot(d, p);
    const float c = dot(p, p) - kEarthRadius * kEarthRadius;
    return b * b - 4.0f * c >= 0.0f;
}

// optical-depth factors from p to space along d
V3 depth_to_space(V3 p, V3 d) {
    const Quadrature& q = quadrature();
    const float len = path_length(p, d);
    V3 acc{0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 8; i++) {
        const V3 s = p + (q.node[i] * len) * d;
        acc = acc + q.weight[i] * density(norm(s) - kEarthRadius);
    }
    return len * acc;
}

float phase_rayleigh(float mu) { return 3.0f / (16.0f * kPi) * (1.0f + mu * mu); }
// Cornette-Shanks phase function
float phase_mie(float mu) {
    const float g2 = kMieG * kMieG;
    return (3.0f * (1.0f - g2) * (1.0f + mu * mu)) /
           (8.0f * kPi * (2.0f + g2) * std::pow(1.0f + g2 - 2.0f * kMieG * mu, 1.5f));
}

V3 xyz_of(const float* spectrum) {
    V3 r{0.0f, 0.0f, 0.0f};
    for (int i = 0; i < kBands; i++) {
        const Band& b = kBandTable[i];
        r = r + spectrum[i] * V3{b.x, b.y, b.z};
    }
    return kBandWidth * r;
}

V3 direction(float elevation, float azimuth) {
    return {std::cos(elevation) * std::cos(azimuth), std::cos(elevation) * std::sin(azimuth), std::sin(elevation)};
}

// in-scattered spectral radiance toward `eye` from direction `view`
void inscatter(V3 view, V3 sun, V3 eye, V3 medium, float* spectrum) {
    const float seg = path_length(eye, view) / kSegments;
    const float mu = dot(view, sun);
    const float pr = phase_rayleigh(mu), pm = phase_mie(mu);
    for (int i = 0; i < kBands; i++) spectrum[i] = 0.0f;
    V3 view_depth{0.0f, 0.0f, 0.0f};
    V3 p = eye + (0.5f * seg) * view;
    for (int k = 0; k < kSegments; k++, p = p + seg * view) {
        const V3 rho = medium * density(norm(p) - kEarthRadius);
        view_depth = view_depth + seg * rho;
        if (hits_ground(p, sun)) continue;  // the sun is below this point's horizon
        const V3 depth = view_depth + medium * depth_to_space(p, sun);
        const float scatter_r = rho.x * pr, scatter_m = rho.y * pm;
        for (int i = 0; i < kBands; i++) {
            const Band& b = kBandTable[i];
            const float tau = depth.x * b.rayleigh + depth.y * (kMieExtinction * kMieScattering) + depth.z * b.ozone;
            spectrum[i] += std::exp(-tau) * (scatter_r * b.rayleigh + scatter_m * kMieScattering) * b.irradiance * seg;
        }
    }
}

}  // namespace

// Rows [y0, y1) of the width x height sky texture (RGBA floats, stride
// floats per texel, alpha untouched): texel (x, y) holds the XYZ sky
// radiance at azimuth 2 pi x / width - pi and elevation
// (pi/2 + pi/(4 height)) (y / height)^2 (more rows toward the horizon);
// the two halves of a row are mirror images about the sun's azimuth 0.
void precompute_texture(float* pixels, int stride, int y0, int y1, int width, int height, float sun_elevation,
                        float altitude, float air_density, float dust_density, float ozone_density) {
    const V3 eye{0.0f, 0.0f, kEarthRadius + altitude};
    const V3 sun = direction(sun_elevation, 0.0f);
    const V3 medium{air_density, dust_density, ozone_density};
    const float lat_step = 0.5f * kPi / height;
    const float lon_step = 2.0f * kPi / width;
    float spectrum[kBands];
    for (int y = y0; y < y1; y++) {
        const float t = (float)y / height;
        const float elevation = (0.5f * kPi + lat_step / 2.0f) * (t * t);
        float* row = pixels + (size_t)y * width * stride;
        for (int x = 0; x < width / 2; x++) {
            inscatter(direction(elevation, lon_step * x - kPi), sun, eye, medium, spectrum);
            const V3 c = xyz_of(spectrum);
            float* a = row + (size_t)x * stride;
            float* b = row + (size_t)(width - 1 - x) * stride;
            a[0] = b[0] = c.x;
            a[1] = b[1] = c.y;
            a[2] = b[2] = c.z;
        }
    }
}

// XYZ radiance of the sun disc's lower and upper edge (clamped to the
// horizon), seen through the atmosphere, per steradian of the disc
void precompute_sun(float sun_elevation, float angular_diameter, float altitude, float air_density,
                    float dust_density, float* bottom, float* top) {
    const float half = angular_diameter / 2.0f;
    const float solid_angle = 2.0f * kPi * (1.0f - std::cos(half));
    const V3 eye{0.0f, 0.0f, kEarthRadius + altitude};
    float spectrum[kBands];
    const float edges[2] = {std::max(sun_elevation - half, 0.0f), std::max(sun_elevation + half, 0.0f)};
    float* outs[2] = {bottom, top};
    for (int e = 0; e < 2; e++) {
        const V3 depth = depth_to_space(eye, direction(edges[e], 0.0f));
        for (int i = 0; i < kBands; i++) {
            const Band& b = kBandTable[i];
            const float tau = b.rayleigh * depth.x * air_density + kMieExtinction * kMieScattering * depth.y * dust_density;
            spectrum[i] = b.irradiance * std::exp(-tau) / solid_angle;
        }
        const V3 c = xyz_of(spectrum);
        outs[e][0] = c.x;
        outs[e][1] = c.y;
        outs[e][2] = c.z;
    }
}

}  // namespace sky
}  // namespace rm

extern "C" {

int rmh_sky_nishita_texture(float* pixels, int width, int height, float sun_elevation, float altitude,
                            float air_density, float dust_density, float ozone_density) {
    if (!pixels || width <= 0 || height <= 0) return -1;
    rm::sky::precompute_texture(pixels, 4, 0, height, width, height, sun_elevation, altitude, air_density,
                                dust_density, ozone_density);
    return 0;
}

int rmh_sky_nishita_sun(float sun_elevation, float angular_diameter, float altitude, float air_density,
                        float dust_density, float* bottom, float* top) {
    if (!bottom || !top) return -1;
    rm::sky::precompute_sun(sun_elevation, angular_diameter, altitude, air_density, dust_density, bottom, top);
    return 0;
}

}  // extern "C"
