// Host BVH8 builder (replaces the Metal acceleration-structure builds of
// host/scene/ShapeBuilder.swift:7-94,300-319 and EntityBuilder.swift:146-165).
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "../device/bvh.h"
#include "../device/common.h"

namespace rm {

struct BuildTri {
    vec3 v0, v1, v2;
    uint32_t inst, prim;
};

struct Bvh8Build {
    std::vector<Bvh8Node> nodes;
    std::vector<BvhTri> tris;
    int depth = 0;           // BVH8 levels
    int bvh2_nodes = 0;
    double build_seconds = 0;
};

// Builds a compressed BVH8 over world-space triangles.  Leaves hold at most
// 3 triangles; `threads` <= 0 uses all hardware threads.
Bvh8Build build_bvh8(const std::vector<BuildTri>& tris, int threads, std::string* error);

}  // namespace rm
