// Allocation-free closest-hit batches: rm_trace_closest_noalloc.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): the caller passes device arrays for the
// rays and the hits, and the intermediate state lives in `scratch`, a
// fixed array of this library in device memory.  Each GPU holds its own
// copy of the array until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for that GPU, taken once
// when the device is created.  Every call on a GPU shares that copy, so
// calls that overlap on one GPU race.
//
// Layout of `scratch` (32-bit words):
//   [0, 64)                 persistent-scheduler counter (word 0), cleared
//                           by a kernel at every call (no memset node)
//   then 10 arrays of kNoAllocMaxRays words: the rays' origin x/y/z, tmin,
//   direction x/y/z, tmax, and the hit barycentrics b0, b1
#include <hip/hip_runtime.h>

#include "../host/core.h"
#include "noalloc_api.h"
#include "trace_api.h"

namespace rm {

constexpr size_t kNoAllocHeadWords = 64;
constexpr size_t kNoAllocWords = kNoAllocHeadWords + 10 * (size_t)kNoAllocMaxRays;

static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

void noalloc_init(rm_device* dev) {
    void* ptr = nullptr;
    RM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    RM_REQUIRE(ptr != nullptr, "no device address for the allocation-free scratch array");
    dev->noalloc_scratch = static_cast<uint32_t*>(ptr);
}

// rays AoS (n x 8) -> SoA columns of the scratch array; clears the counter
__global__ void noalloc_unpack_kernel(const float* __restrict__ rays, uint32_t n, uint32_t* __restrict__ s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) s[0] = 0u;
    if (i >= n) return;
    float* cols = reinterpret_cast<float*>(s + kNoAllocHeadWords);
    const float4 a = reinterpret_cast<const float4*>(rays)[2 * (size_t)i];
    const float4 b = reinterpret_cast<const float4*>(rays)[2 * (size_t)i + 1];
    cols[0 * (size_t)kNoAllocMaxRays + i] = a.x;
    cols[1 * (size_t)kNoAllocMaxRays + i] = a.y;
    cols[2 * (size_t)kNoAllocMaxRays + i] = a.z;
    cols[3 * (size_t)kNoAllocMaxRays + i] = a.w;
    cols[4 * (size_t)kNoAllocMaxRays + i] = b.x;
    cols[5 * (size_t)kNoAllocMaxRays + i] = b.y;
    cols[6 * (size_t)kNoAllocMaxRays + i] = b.z;
    cols[7 * (size_t)kNoAllocMaxRays + i] = b.w;
}

// barycentric columns -> the caller's interleaved pairs
__global__ void noalloc_pack_kernel(const uint32_t* __restrict__ s, uint32_t n, float* __restrict__ bary) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* cols = reinterpret_cast<const float*>(s + kNoAllocHeadWords);
    reinterpret_cast<float2*>(bary)[i] =
        make_float2(cols[8 * (size_t)kNoAllocMaxRays + i], cols[9 * (size_t)kNoAllocMaxRays + i]);
}

void trace_closest_noalloc(rm_scene* sc, const float* rays, uint32_t n, float* t, uint32_t* inst, uint32_t* prim,
                           float* bary, hipStream_t stream) {
    RM_REQUIRE(sc->accel_built, "acceleration structure not built");
    RM_REQUIRE(n <= kNoAllocMaxRays, "rm_trace_closest_noalloc takes at most " + std::to_string(kNoAllocMaxRays) +
                                         " rays per call (got " + std::to_string(n) + ")");
    RM_REQUIRE((reinterpret_cast<uintptr_t>(rays) & 15u) == 0 && (reinterpret_cast<uintptr_t>(bary) & 7u) == 0,
               "rays must be 16-byte and bary 8-byte aligned device arrays");
    uint32_t* s = sc->dev->noalloc_scratch;
    RM_REQUIRE(s != nullptr, "allocation-free scratch array not set up on this device");
    if (n == 0) return;
    hipStream_t st = stream ? stream : sc->dev->stream;
    const uint32_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(noalloc_unpack_kernel, dim3(blocks), dim3(256), 0, st, rays, n, s);
    RM_HIP(hipGetLastError());
    float* cols = reinterpret_cast<float*>(s + kNoAllocHeadWords);
    ClosestArgs a{};
    a.nodes = sc->d_nodes.as<Bvh8Node>();
    a.tlas = sc->two_level ? sc->d_tlas.as<TlasInstance>() : nullptr;
    a.root2l = sc->root2l;
    a.tris = sc->d_tris.as<BvhTri>();
    a.count = n;
    a.fetch = s;
    a.paths.ox = cols;
    a.paths.oy = cols + (size_t)kNoAllocMaxRays;
    a.paths.oz = cols + 2 * (size_t)kNoAllocMaxRays;
    a.tmin_arr = cols + 3 * (size_t)kNoAllocMaxRays;
    a.paths.dx = cols + 4 * (size_t)kNoAllocMaxRays;
    a.paths.dy = cols + 5 * (size_t)kNoAllocMaxRays;
    a.paths.dz = cols + 6 * (size_t)kNoAllocMaxRays;
    a.tmax_arr = cols + 7 * (size_t)kNoAllocMaxRays;
    a.hits.t = t;
    a.hits.inst = inst;
    a.hits.prim = prim;
    a.hits.b0 = cols + 8 * (size_t)kNoAllocMaxRays;
    a.hits.b1 = cols + 9 * (size_t)kNoAllocMaxRays;
    a.spill = sc->d_spill.as<uint2>();
    a.counters = nullptr;
    const int tblocks = sc->dev->props.multiProcessorCount * sc->dev->trace_blocks_per_cu;
    RM_HIP(launch_trace_closest(a, tblocks, false, st));
    hipLaunchKernelGGL(noalloc_pack_kernel, dim3(blocks), dim3(256), 0, st, s, n, bary);
    RM_HIP(hipGetLastError());
}

}  // namespace rm
