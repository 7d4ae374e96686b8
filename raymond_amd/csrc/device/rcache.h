// Learned radiance cache: shared definitions of the hash-grid encoding, the
// MLP shape, the per-vertex input features and the device-side record queues.
//
// Not a reference component.  The method follows the published neural
// radiance cache (Müller, Rousselle, Novák, Keller, "Real-time Neural Radiance
// Caching for Path Tracing", SIGGRAPH 2021) with the multiresolution hash
// encoding of Müller, Evans, Schied, Keller (SIGGRAPH 2022):
//   * encoding: kRcLevels levels x 2 features, 2^log2_table entries per level,
//     resolutions geometric from 16 to 2048; a level whose (res+1)^3 vertices
//     fit the table is indexed densely, finer levels through the spatial hash
//     x*1 ^ y*2654435761 ^ z*805459861 (Teschner et al. 2003);
//   * MLP: 48 inputs (32 encoding + 9 vertex features + a constant 1, zero
//     padded) -> 64 -> 64 -> 3, ReLU on the hidden layers, linear output, no
//     biases; FP16 operands, FP32 accumulation;
//   * loss: relative L2, sum_c (y_c - t_c)^2 / (y_c^2 + 0.01) with y_c held
//     constant in the denominator.
#pragma once
#include "common.h"
#include "elem.h"

namespace rm {

constexpr int kRcLevels = 16;
constexpr int kRcFeaturesPerLevel = 2;
constexpr int kRcEncoded = kRcLevels * kRcFeaturesPerLevel;  // 32
constexpr int kRcExtra = 9;                                  // vertex features after the encoding
constexpr int kRcInputs = 48;                                // 32 + 9 + constant 1, padded to 16
constexpr int kRcHidden = 64;
constexpr int kRcOutputs = 3;
constexpr int kRcFeatureFloats = 3 + kRcExtra;  // per-record features: position in [0,1]^3 + extras

// MLP parameter block, row-major [in][out] per layer; W3 keeps 32 output
// columns so every layer has the MFMA tile width (columns 3..31 stay zero).
constexpr int kRcW1 = kRcInputs * kRcHidden;   // 3072
constexpr int kRcW2 = kRcHidden * kRcHidden;   // 4096
constexpr int kRcW3Cols = 32;
constexpr int kRcW3 = kRcHidden * kRcW3Cols;   // 2048
constexpr int kRcMlpParams = kRcW1 + kRcW2 + kRcW3;

struct RcGrid {
    uint32_t log2_table;           // entries per level = 1 << log2_table
    uint32_t res[kRcLevels];       // grid cells per axis
    uint32_t dense[kRcLevels];     // 1 if (res+1)^3 <= table size
};

// Per-level vertex index of grid point (x, y, z), each in [0, res].
RM_HD uint32_t rc_index(const RcGrid& g, int level, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t mask = (1u << g.log2_table) - 1u;
    if (g.dense[level]) {
        const uint32_t n = g.res[level] + 1u;
        return x + n * (y + n * z);
    }
    return ((x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u)) & mask;
}

// Cell and fractional position of p (in [0,1]) at a level of `res` cells.
RM_HD void rc_cell(float p, uint32_t res, uint32_t& cell, float& frac) {
    const float x = p * (float)res;
    float c = floor_(x);
    c = min_(max_(c, 0.0f), (float)(res - 1u));
    cell = (uint32_t)c;
    frac = x - c;
}

// atan2 for network inputs, in float: octant reduction to [0, 1] and the
// odd polynomial of Abramowitz and Stegun 4.4.49 (error below 2e-8 rad in
// exact arithmetic).  The features only feed the network, so they need none
// of the correctly rounded elementary functions the shading code uses (those
// run in double precision and cost about ten times as much).
RM_HD float rc_atan2(float y, float x) {
    const float ax = x < 0.0f ? -x : x, ay = y < 0.0f ? -y : y;
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    const float t = mx > 0.0f ? mn / mx : 0.0f;
    const float t2 = t * t;
    float p = 0.0028662257f;
    p = p * t2 - 0.0161657367f;
    p = p * t2 + 0.0429096138f;
    p = p * t2 - 0.0752896400f;
    p = p * t2 + 0.1065626393f;
    p = p * t2 - 0.1420889944f;
    p = p * t2 + 0.1999355085f;
    p = p * t2 - 0.3333314528f;
    p = p * t2 + 1.0f;
    float a = p * t;
    a = ay > ax ? 1.5707963268f - a : a;
    a = x < 0.0f ? 3.1415926536f - a : a;
    return y < 0.0f ? -a : a;
}
RM_HD float rc_acos(float z) {
    z = clamp_(z, -1.0f, 1.0f);
    return rc_atan2(sqrt_((1.0f - z) * (1.0f + z)), z);
}

// The 9 vertex features: outgoing direction and normal in spherical
// coordinates (scaled to [0,1]), roughness, albedo and the specular-lobe
// probability.
RM_HD void rc_vertex_features(vec3 wo, vec3 n, float roughness, vec3 albedo, float spec, float* f) {
    f[0] = rc_acos(wo.z) * kInvPi;
    f[1] = rc_atan2(wo.y, wo.x) * (0.5f * kInvPi) + 0.5f;
    f[2] = rc_acos(n.z) * kInvPi;
    f[3] = rc_atan2(n.y, n.x) * (0.5f * kInvPi) + 0.5f;
    f[4] = roughness;
    f[5] = saturate(albedo.x);
    f[6] = saturate(albedo.y);
    f[7] = saturate(albedo.z);
    f[8] = spec;
}

// Queries emitted by the shading kernel: the path ends at this vertex and
// `throughput x cache(features)` is added to `pixel`.  One 64-byte row per
// query: [0] pixel (bits), [1..3] throughput, [4..15] vertex features, so a
// lookup read in any order touches one cache line.
constexpr int kRcQueryFloats = 16;
struct RcQueryQueue {
    float* rows;  // kRcQueryFloats x capacity, row-major
};

// Vertices of training paths: features, the path throughput on arrival and
// the path's radiance sum after this vertex's emission.  The target is
// (radiance at path end - r) / t.
struct RcTrainQueue {
    uint32_t* slot;
    float *tx, *ty, *tz;
    float *rx, *ry, *rz;
    float* feat;  // kRcFeatureFloats x capacity
};

// What the bounce loop needs to know about the cache.
struct CacheView {
    uint32_t enabled;        // query at the first vertex after a diffuse bounce
    uint32_t train_shift;    // a camera path trains with probability 2^-train_shift
    uint32_t query_capacity;
    uint32_t train_capacity; // vertex records
    uint32_t slot_capacity;  // training paths
    vec3 bounds_min;
    vec3 bounds_inv_extent;
    uint32_t* slots;         // training slot per entry of the current path queue
    uint32_t* slots_next;    // ... of the next path queue
    float* slot_radiance;    // RGB per training path
    RcQueryQueue queries;
    RcTrainQueue train;
};

// Training paths.  Samples g = 0, 1, ... of a bounce loop form groups of
// 2^shift consecutive samples (consecutive ray indices of one frame); one
// sample per group, at a hashed position, is a training path and gets slot
// g >> shift.  The choice is independent of the path's own random stream.
constexpr uint32_t kRcNoSlot = 0xFFFFFFFFu;
constexpr uint32_t kRcRecorded = 0x80000000u;  // the path has recorded its training vertex

RM_HD uint32_t rc_training_slot(uint32_t g, uint32_t seed, uint32_t shift) {
    const uint32_t group = g >> shift;
    uint32_t h = group * 0x9E3779B1u ^ seed * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    const uint32_t mask = (1u << shift) - 1u;
    return (g & mask) == (h & mask) ? group : kRcNoSlot;
}

}  // namespace rm
