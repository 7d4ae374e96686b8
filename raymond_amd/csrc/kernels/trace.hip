// Traversal kernels for gfx950: closest hit (optionally fused with camera-ray
// generation) and any hit (fused with shadow-ray accumulation).
//
// Counterparts: generateRays (device/kernels/shading/generate.metal:72-192),
// raytrace / raytraceAny (device/kernels/shading/raytrace.metal:11-85) and
// handleShadowRays (device/kernels/shading/shadow.metal:6-29).
//
// Design:
//  * persistent waves: each 64-lane workgroup repeatedly claims 64 rays with
//    one atomic on a per-depth fetch counter until the queue (whose length
//    lives in device memory) is exhausted, so launches need no host
//    round-trip and lanes of short rays are refilled;
//  * compressed BVH8 (80-byte nodes, five 16-byte loads, 8-bit child boxes)
//    with ordered internal-child visits by ray octant;
//  * traversal stack: the bottom kLdsStack entries per lane in LDS (column
//    per lane), deeper entries spill to a per-thread global array.
#include <hip/hip_runtime.h>
#include "../device/common.h"
#include "../device/rng.h"
#include "../device/isect.h"
#include "../device/bvh.h"
#include "../device/queues.h"
#include "../device/camera.h"
#include "../device/scene.h"
#include "trace_api.h"

namespace rm {

constexpr int kLdsStack = 16;

__device__ inline uint32_t byte_of(uint32_t lo, uint32_t hi, int c) {
    const uint32_t w = c < 4 ? lo : hi;
    return (w >> (8 * (c & 3))) & 0xffu;
}

struct StackRef {
    uint2* lds;      // column base: lds[sp * 64]
    uint2* spill;    // spill[(sp - kLdsStack) * stride]
    uint32_t stride;
    __device__ void store(int sp, uint2 v) const {
        if (sp < kLdsStack) lds[sp * kTraceBlock] = v;
        else spill[(size_t)(sp - kLdsStack) * stride] = v;
    }
    __device__ uint2 load(int sp) const {
        return sp < kLdsStack ? lds[sp * kTraceBlock] : spill[(size_t)(sp - kLdsStack) * stride];
    }
};

struct TraceResult {
    float t;
    uint32_t inst, prim;
    float b0, b1;
    uint32_t nodes, tris;
};

// Returns true on a hit. ANY: stop at the first hit in (tmin, tmax).
template <bool ANY, bool STATS>
__device__ bool traverse(const Bvh8Node* __restrict__ nodes, const BvhTri* __restrict__ tris, vec3 o, vec3 d,
                         float tmin, float tmax, const StackRef& stk, TraceResult& res) {
    const vec3 inv = safe_inv_dir(d);
    const RayShear sh = make_shear(d);
    const int oct = dir_octant(d);
    uint2 cur = make_uint2(0u, (1u << (24 + oct)) | 1u);
    int sp = 0;
    float best_t = tmax;
    uint32_t best_inst = 0xffffffffu, best_prim = 0xffffffffu;
    float bb0 = 0.0f, bb1 = 0.0f;
    bool found = false;
    uint32_t n_nodes = 0, n_tris = 0;

    while (true) {
        uint2 tri = make_uint2(0u, 0u);
        if (cur.y & 0xff000000u) {
            const uint32_t hits = cur.y >> 24;
            const int k = __builtin_ctz(hits);
            cur.y &= ~(1u << (24 + k));
            const int slot = k ^ oct;
            const uint32_t node_idx = cur.x + __builtin_popcount(cur.y & 0xffu & ((1u << slot) - 1u));
            if (cur.y & 0xff000000u) {
                stk.store(sp, cur);
                sp++;
            }
            const uint4* np = reinterpret_cast<const uint4*>(nodes + node_idx);
            const uint4 q0 = np[0], q1 = np[1], q2 = np[2], q3 = np[3], q4 = np[4];
            if (STATS) n_nodes++;
            const float px = __uint_as_float(q0.x), py = __uint_as_float(q0.y), pz = __uint_as_float(q0.z);
            const float sx = exp_scale((uint8_t)(q0.w & 0xffu));
            const float sy = exp_scale((uint8_t)((q0.w >> 8) & 0xffu));
            const float sz = exp_scale((uint8_t)((q0.w >> 16) & 0xffu));
            const uint32_t imask = q0.w >> 24;
            uint32_t ihits = 0, thits = 0;
#pragma unroll
            for (int c = 0; c < 8; c++) {
                const uint32_t meta = byte_of(q1.z, q1.w, c);
                if (meta == 0) continue;
                const vec3 lo(decode_q(px, (uint8_t)byte_of(q2.x, q2.y, c), sx),
                              decode_q(py, (uint8_t)byte_of(q2.z, q2.w, c), sy),
                              decode_q(pz, (uint8_t)byte_of(q3.x, q3.y, c), sz));
                const vec3 hi(decode_q(px, (uint8_t)byte_of(q3.z, q3.w, c), sx),
                              decode_q(py, (uint8_t)byte_of(q4.x, q4.y, c), sy),
                              decode_q(pz, (uint8_t)byte_of(q4.z, q4.w, c), sz));
                float tn;
                if (box_hit(lo, hi, o, inv, tmin, best_t, tn)) {
                    if ((imask >> c) & 1u) ihits |= 1u << (c ^ oct);
                    else thits |= (meta >> 5) << (meta & 31u);
                }
            }
            cur = make_uint2(q1.x, (ihits << 24) | imask);
            tri = make_uint2(q1.y, thits);
        } else {
            if (sp == 0) break;
            sp--;
            cur = stk.load(sp);
            continue;
        }
        while (tri.y) {
            const int b = __builtin_ctz(tri.y);
            tri.y &= tri.y - 1u;
            const uint4* tp = reinterpret_cast<const uint4*>(tris + tri.x + b);
            const uint4 a = tp[0], bq = tp[1], cq = tp[2];
            if (STATS) n_tris++;
            const vec3 v0(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z));
            const vec3 v1(__uint_as_float(bq.x), __uint_as_float(bq.y), __uint_as_float(bq.z));
            const vec3 v2(__uint_as_float(cq.x), __uint_as_float(cq.y), __uint_as_float(cq.z));
            float t, b0, b1;
            if (intersect_triangle(o, sh, v0, v1, v2, tmin, tmax, t, b0, b1) &&
                hit_before(t, a.w, bq.w, best_t, best_inst, best_prim)) {
                best_t = t;
                best_inst = a.w;
                best_prim = bq.w;
                bb0 = b0;
                bb1 = b1;
                found = true;
                if (ANY) {
                    res.nodes = n_nodes;
                    res.tris = n_tris;
                    res.t = t;
                    return true;
                }
            }
        }
    }
    res.t = found ? best_t : -1.0f;
    res.inst = best_inst;
    res.prim = best_prim;
    res.b0 = bb0;
    res.b1 = bb1;
    res.nodes = n_nodes;
    res.tris = n_tris;
    return found;
}

__device__ inline uint32_t wave_fetch(uint32_t* counter) {
    uint32_t base = 0;
    if (threadIdx.x == 0) base = atomicAdd(counter, (uint32_t)kTraceBlock);
    return __shfl(base, 0);
}

// Closest-hit over a queue; mode kTraceCamera also generates the camera rays
// and writes the initial path state.
template <bool STATS>
__global__ __launch_bounds__(kTraceBlock) void trace_closest_kernel(ClosestArgs a) {
    __shared__ uint2 lds[kLdsStack * kTraceBlock];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * kTraceBlock + lane;
    StackRef stk{lds + lane, a.spill + gtid, gridDim.x * kTraceBlock};
    const uint32_t count = a.count_ptr ? *a.count_ptr : a.count;
    unsigned long long sn = 0, st = 0;
    while (true) {
        const uint32_t base = wave_fetch(a.fetch);
        if (base >= count) break;
        const uint32_t i = base + lane;
        if (i >= count) continue;
        vec3 o, d;
        float tmin, tmax;
        if (a.mode == kTraceCamera) {
            const uint32_t wh = a.width * a.height;
            const uint32_t s = i / wh;
            const uint32_t ray_index = i - s * wh;
            uint32_t x, y;
            block_linear_to_pixel(ray_index, a.width, a.height, x, y);
            Prng prng(a.random_seed + s, ray_index);
            camera_ray(a.camera, x, y, a.width, a.height, prng, o, d);
            tmin = a.camera.near_clip;
            tmax = a.camera.far_clip;
            PathQueue& q = a.paths;
            q.ox[i] = o.x; q.oy[i] = o.y; q.oz[i] = o.z;
            q.dx[i] = d.x; q.dy[i] = d.y; q.dz[i] = d.z;
            q.wx[i] = 1.0f; q.wy[i] = 1.0f; q.wz[i] = 1.0f;
            q.pdf[i] = kInf;
            q.seed[i] = prng.seed;
            q.meta[i] = pack_meta(prng.index, RayCamera, 0);
            q.pixel[i] = a.image_offset + y * a.width + x;
        } else {
            const PathQueue& q = a.paths;
            o = vec3(q.ox[i], q.oy[i], q.oz[i]);
            d = vec3(q.dx[i], q.dy[i], q.dz[i]);
            tmin = a.tmin_arr ? a.tmin_arr[i] : kEps;
            tmax = a.tmax_arr ? a.tmax_arr[i] : kInf;
        }
        TraceResult r;
        traverse<false, STATS>(a.nodes, a.tris, o, d, tmin, tmax, stk, r);
        a.hits.t[i] = r.t;
        a.hits.prim[i] = r.prim;
        a.hits.inst[i] = r.inst;
        a.hits.b0[i] = r.b0;
        a.hits.b1[i] = r.b1;
        if (STATS) { sn += r.nodes; st += r.tris; }
    }
    if (STATS) {
        atomicAdd(&a.counters->nodes_closest, sn);
        atomicAdd(&a.counters->tris_closest, st);
    }
}

// Any-hit over the shadow queue.  With accumulate, an unoccluded ray adds
// (weight, 1) to its pixel (shadow.metal:6-29); otherwise the occlusion flag
// is written to `occluded`.
template <bool STATS>
__global__ __launch_bounds__(kTraceBlock) void trace_any_kernel(AnyArgs a) {
    __shared__ uint2 lds[kLdsStack * kTraceBlock];
    const uint32_t lane = threadIdx.x;
    const uint32_t gtid = blockIdx.x * kTraceBlock + lane;
    StackRef stk{lds + lane, a.spill + gtid, gridDim.x * kTraceBlock};
    const uint32_t count = a.count_ptr ? *a.count_ptr : a.count;
    unsigned long long sn = 0, st = 0;
    while (true) {
        const uint32_t base = wave_fetch(a.fetch);
        if (base >= count) break;
        const uint32_t i = base + lane;
        if (i >= count) continue;
        const ShadowQueue& q = a.shadows;
        const vec3 o(q.ox[i], q.oy[i], q.oz[i]);
        const vec3 d(q.dx[i], q.dy[i], q.dz[i]);
        const float tmin = a.tmin_arr ? a.tmin_arr[i] : kEps;
        const float tmax = q.tmax[i];
        TraceResult r;
        const bool hit = traverse<true, STATS>(a.nodes, a.tris, o, d, tmin, tmax, stk, r);
        if (a.image) {
            if (!hit) {
                float* px = a.image + 4u * (size_t)q.pixel[i];
                atomicAdd(px + 0, q.wx[i]);
                atomicAdd(px + 1, q.wy[i]);
                atomicAdd(px + 2, q.wz[i]);
                atomicAdd(px + 3, 1.0f);
            }
        } else {
            a.occluded[i] = hit ? 1u : 0u;
        }
        if (STATS) { sn += r.nodes; st += r.tris; }
    }
    if (STATS) {
        atomicAdd(&a.counters->nodes_any, sn);
        atomicAdd(&a.counters->tris_any, st);
    }
}

}  // namespace rm

// ---------------------------------------------------------------------------
// Host launchers

namespace rm {

hipError_t launch_trace_closest(const ClosestArgs& a, int blocks, bool stats, hipStream_t stream) {
    if (stats) hipLaunchKernelGGL(trace_closest_kernel<true>, dim3(blocks), dim3(kTraceBlock), 0, stream, a);
    else hipLaunchKernelGGL(trace_closest_kernel<false>, dim3(blocks), dim3(kTraceBlock), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_trace_any(const AnyArgs& a, int blocks, bool stats, hipStream_t stream) {
    if (stats) hipLaunchKernelGGL(trace_any_kernel<true>, dim3(blocks), dim3(kTraceBlock), 0, stream, a);
    else hipLaunchKernelGGL(trace_any_kernel<false>, dim3(blocks), dim3(kTraceBlock), 0, stream, a);
    return hipGetLastError();
}

int trace_blocks_per_cu(bool any) {
    int n = 0;
    if (any) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_any_kernel<false>, kTraceBlock, 0);
    else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, trace_closest_kernel<false>, kTraceBlock, 0);
    return n;
}

}  // namespace rm
