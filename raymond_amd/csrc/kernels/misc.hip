// Small ahead-of-time kernels: counter reset, envmap mip reduction and pdf
// normalisation, image resolve/tonemap, FP64 accumulation.
#include <hip/hip_runtime.h>
#include "../device/common.h"
#include "../device/queues.h"
#include "misc_api.h"

namespace rm {

__global__ void clear_counters_kernel(Counters* c) {
    uint32_t* w = reinterpret_cast<uint32_t*>(c);
    const uint32_t n = (uint32_t)(sizeof(Counters) / 4);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = 0u;
}

// One quad-tree level: sums the 4 children of each texel in the next finer
// level, normalises them and writes the sum (device/kernels/envmap/reduce.metal:3-24).
__global__ void envmap_reduce_kernel(float* level, int res) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (uint32_t)(res * res)) return;
    const uint32_t x = tid % (uint32_t)res, y = tid / (uint32_t)res;
    const uint32_t input = y * (uint32_t)res + x;
    const uint32_t len = (uint32_t)(res * res);
    float* child = level + len + 4u * input;
    float sum = 0.0f;
    for (int i = 0; i < 4; ++i) sum += child[i];
    for (int i = 0; i < 4; ++i) child[i] /= sum;
    if (res == 1) {
        level[0] = sum;
        return;
    }
    const uint32_t quad = (y / 2u) * ((uint32_t)res / 2u) + (x / 2u);
    level[4u * quad + (x & 1u) + 2u * (y & 1u)] = sum;
}

// pdf = value * N / (4 pi * total) (device/kernels/envmap/normalize.metal:4-11)
__global__ void envmap_normalize_kernel(const float* total, float* pdfs, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pdfs[i] = pdfs[i] * (((float)n * (1.0f / (4.0f * kPi))) / *total);
}

__device__ inline vec3 hable_partial(vec3 x) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    auto f = [&](float v) { return ((v * (A * v + C * B) + D * E) / (v * (A * v + B) + D * F)) - E / F; };
    return vec3(f(x.x), f(x.y), f(x.z));
}

// (device/kernels/utils/normalizeImage.metal:11-59)
__global__ void resolve_kernel(const float* img, float* out, uint32_t n, float exposure, float divisor, int tonemap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vec3 c = exposure * vec3(img[4 * i], img[4 * i + 1], img[4 * i + 2]);
    c = c / divisor;
    if (tonemap == 1) {
        const vec3 curr = hable_partial(c * 2.0f);
        const vec3 white = 1.0f / hable_partial(vec3(11.2f));
        c = curr * white;
    } else if (tonemap == 2) {
        const float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, e = 0.14f;
        auto f = [&](float x) { return saturate((x * (a * x + b)) / (x * (cc * x + d) + e)); };
        c = vec3(f(c.x), f(c.y), f(c.z));
    }
    out[4 * i + 0] = c.x;
    out[4 * i + 1] = c.y;
    out[4 * i + 2] = c.z;
    out[4 * i + 3] = 1.0f;
}

// FP64 accumulation of one launch; clears the FP32 launch image.
__global__ void accum_f64_kernel(float* img, double* acc, uint32_t n4) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n4) return;
    acc[i] += (double)img[i];
    img[i] = 0.0f;
}

hipError_t launch_clear_counters(Counters* c, hipStream_t s) {
    hipLaunchKernelGGL(clear_counters_kernel, dim3(64), dim3(256), 0, s, c);
    return hipGetLastError();
}
// Streaming copy of n 16-byte words: the reference point for hardware byte
// counters (it reads and writes exactly 16 n bytes).
__global__ void copy_probe_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy_probe(const void* src, void* dst, size_t bytes, hipStream_t s) {
    hipLaunchKernelGGL(copy_probe_kernel, dim3(4096), dim3(256), 0, s, reinterpret_cast<const uint4*>(src),
                       reinterpret_cast<uint4*>(dst), bytes / 16);
    return hipGetLastError();
}

// Moves a captured bounce-loop graph on to its next chunk of iterations.
__global__ void advance_iter_kernel(Counters* c, uint32_t n) {
    if (threadIdx.x == 0) c->iter_base += n;
}

hipError_t launch_advance_iter(Counters* c, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(advance_iter_kernel, dim3(1), dim3(64), 0, s, c, n);
    return hipGetLastError();
}

hipError_t launch_envmap_reduce(float* level, int res, hipStream_t s) {
    const uint32_t n = (uint32_t)(res * res);
    hipLaunchKernelGGL(envmap_reduce_kernel, dim3((n + 255) / 256), dim3(256), 0, s, level, res);
    return hipGetLastError();
}
hipError_t launch_envmap_normalize(const float* total, float* pdfs, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(envmap_normalize_kernel, dim3((n + 255) / 256), dim3(256), 0, s, total, pdfs, n);
    return hipGetLastError();
}
hipError_t launch_resolve(const float* img, float* out, uint32_t n, float exposure, float divisor, int tonemap,
                          hipStream_t s) {
    hipLaunchKernelGGL(resolve_kernel, dim3((n + 255) / 256), dim3(256), 0, s, img, out, n, exposure, divisor, tonemap);
    return hipGetLastError();
}
hipError_t launch_accum_f64(float* img, double* acc, uint32_t n4, hipStream_t s) {
    hipLaunchKernelGGL(accum_f64_kernel, dim3((n4 + 255) / 256), dim3(256), 0, s, img, acc, n4);
    return hipGetLastError();
}

}  // namespace rm
