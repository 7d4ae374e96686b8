// Device-visible scene layout: the counterpart of the reference's Context
// argument buffer (device/Context.hpp:15-26), PerInstanceData
// (bridge/PerInstanceData.hpp:6-20), Camera (bridge/Camera.hpp:5-11) and the
// light records (bridge/lights/*.hpp).  Plain structs of device pointers;
// one copy lives in device memory per scene and kernels take its address.
#pragma once
#include "common.h"

namespace rm {

struct PerInstance {
    uint32_t vertex_offset;
    uint32_t face_offset;
    uint32_t light_face_offset;
    uint32_t light_face_count;
    uint32_t light_index;
    uint32_t visibility;  // RayFlags mask
    vec3 bounds_min;
    vec3 bounds_size;
    mat4 point_transform;
    mat3 normal_transform;  // inverse-transpose of the inner 3x3
};

struct Camera {
    mat4 transform;
    float near_clip;
    float far_clip;
    float focal_length;  // f_mm / (film_width / 2)
    vec2 shift;
};

struct LightInfo {
    uint32_t shader_index;
    uint32_t casts_shadows;
    uint32_t uses_mis;
};

struct AreaLight {
    LightInfo info;
    vec4 t0, t1, t2;  // float3x4 columns: point = (dot(p,t0), dot(p,t1), dot(p,t2))
    vec3 color;
    uint32_t is_circular;
};
struct PointLight {
    LightInfo info;
    vec3 location;
    float radius;
    vec3 color;
};
struct SunLight {
    LightInfo info;
    vec3 direction;
    float cos_angle;
    vec3 color;
};
struct SpotLight {
    LightInfo info;
    vec3 location;
    vec3 direction;
    float radius;
    vec3 color;
    float spot_size;
    float spot_blend;
};
struct ShapeLight {
    uint32_t instance_index;
    float emissive_area;
};

// Importance-sampled environment (device/lights/WorldLight.hpp:14-62)
struct WorldLight {
    uint32_t shader_index;
    int resolution;         // power of two
    const float* pdfs;      // resolution^2
    const float* mipmap;    // sum_k 4^k, quad-tree of normalised child probabilities
};

struct LightsView {
    int num_total;
    int num_area;
    int num_point;
    int num_sun;
    int num_spot;
    int num_shape;
    const float* light_faces;  // per shape light, face-area CDF
    WorldLight world;
    const AreaLight* area;
    const PointLight* point;
    const SunLight* sun;
    const SpotLight* spot;
    const ShapeLight* shape;
};

// An image of the material library (the reference's ctx.textures[i]):
// width x height texels of `channels` floats (1: R, 4: RGBA), row 0 first.
struct Texture {
    const float* pixels;
    int width, height, channels;
};

struct SceneView {
    const float* vertices;     // packed float3, object space
    const uint32_t* indices;   // 3 per face, shape-local vertex indices
    const float* normals;      // packed float3
    const float* texcoords;    // float2
    const PerInstance* instances;
    const uint16_t* materials; // per face
    Camera camera;
    LightsView lights;
    const Texture* textures;   // slots of the material library's images
    uint32_t num_textures;

    RM_HD vec3 vertex(uint32_t i) const { return vec3(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]); }
    RM_HD vec3 normal(uint32_t i) const { return vec3(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]); }
    RM_HD vec2 texcoord(uint32_t i) const { return vec2(texcoords[2 * i], texcoords[2 * i + 1]); }
};

// Closest-hit record (bridge/Ray.hpp:41-46).  coords = barycentric weights of
// vertex 0 and vertex 1; vertex 2 is the interpolation base.  A miss has
// distance < 0 (consumers test distance <= 0, device/kernels/shading/intersection.metal:75).
struct Hit {
    float distance;
    uint32_t prim;
    uint32_t inst;
    vec2 coords;
};

}  // namespace rm
