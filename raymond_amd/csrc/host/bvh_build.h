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

// Builds a compressed BVH8 over triangles.  Leaves hold at most `max_leaf`
// (1..3) triangles; `threads` <= 0 uses all hardware threads.  A primitive
// that is a box (an instance of a two-level tree) is passed as the
// degenerate triangle (lo, hi, hi), whose bounds are the box.
Bvh8Build build_bvh8(const std::vector<BuildTri>& tris, int threads, std::string* error, int max_leaf = 3);

}  // namespace rm
