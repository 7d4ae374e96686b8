// Counting sort of the shading queue by material id.
//
// The shading kernel evaluates a material graph per hit; with many materials
// a 64-lane wave of unsorted hits runs one graph body per distinct material
// it holds.  Grouping the queue by material first gives each wave one or two
// bodies.  Two kernels per bounce-loop iteration, both sized by the queue
// length the iteration's counters hold:
//   keys    : key per entry (face material, or the miss bin) and the global
//             bin counts, aggregated per workgroup in LDS;
//   scatter : exclusive prefix of the bins, one atomic per (workgroup, bin)
//             to claim a range, LDS ranks inside it, order[pos] = entry.
// The permutation is all that moves; paths and hits stay where they are and
// the shading kernel reads them through `order`.
#include "sort_api.h"

namespace rm {

static_assert(kSortBlock == kSortBins, "the scatter scan assigns one bin per thread");

__global__ __launch_bounds__(kSortBlock) void sort_keys_kernel(SortArgs a) {
    __shared__ uint32_t hist[kSortBins];
    const uint32_t t = threadIdx.x;
    hist[t] = 0u;
    __syncthreads();
    const uint32_t iter = iteration_of(a.iter_dev, a.iter);
    const uint32_t count = a.counters->it[iter].queue;
    const uint32_t stride = gridDim.x * kSortBlock;
    for (uint32_t i = blockIdx.x * kSortBlock + t; i < count; i += stride) {
        uint32_t key = kSortBins - 1;  // miss: the world shader
        if (a.hits.t[i] > 0.0f) {
            const uint32_t face = a.instances[a.hits.inst[i]].face_offset + a.hits.prim[i];
            const uint32_t m = (uint32_t)a.materials[face] >> a.key_shift;
            key = m < (uint32_t)(kSortBins - 2) ? m : (uint32_t)(kSortBins - 2);
        }
        a.keys[i] = key;
        atomicAdd(&hist[key], 1u);
    }
    __syncthreads();
    if (hist[t]) atomicAdd(&a.counters->sort_bins[iter & 1u][t], hist[t]);
}

__global__ __launch_bounds__(kSortBlock) void sort_scatter_kernel(SortArgs a) {
    __shared__ uint32_t scan[kSortBins];
    __shared__ uint32_t start[kSortBins];
    __shared__ uint32_t cnt[kSortBins];
    const uint32_t t = threadIdx.x;
    const uint32_t iter = iteration_of(a.iter_dev, a.iter);
    const uint32_t par = iter & 1u;

    // exclusive prefix over the bins (Hillis-Steele in LDS)
    const uint32_t own = a.counters->sort_bins[par][t];
    scan[t] = own;
    cnt[t] = 0u;
    __syncthreads();
    for (uint32_t off = 1; off < (uint32_t)kSortBins; off <<= 1) {
        const uint32_t add = t >= off ? scan[t - off] : 0u;
        __syncthreads();
        scan[t] += add;
        __syncthreads();
    }
    start[t] = scan[t] - own;

    // this workgroup's count per bin, then one claim per non-empty bin
    const uint32_t count = a.counters->it[iter].queue;
    const uint32_t stride = gridDim.x * kSortBlock;
    for (uint32_t i = blockIdx.x * kSortBlock + t; i < count; i += stride) atomicAdd(&cnt[a.keys[i]], 1u);
    __syncthreads();
    const uint32_t mine = cnt[t];
    if (mine) start[t] += atomicAdd(&a.counters->sort_cursor[par][t], mine);
    cnt[t] = 0u;
    __syncthreads();

    for (uint32_t i = blockIdx.x * kSortBlock + t; i < count; i += stride) {
        const uint32_t k = a.keys[i];
        a.order[start[k] + atomicAdd(&cnt[k], 1u)] = i;
    }

    // the other parity was last used by iteration iter-1, which has finished
    if (blockIdx.x == 0) {
        a.counters->sort_bins[par ^ 1u][t] = 0u;
        a.counters->sort_cursor[par ^ 1u][t] = 0u;
    }
}

hipError_t launch_sort_keys(const SortArgs& a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(sort_keys_kernel, dim3(blocks), dim3(kSortBlock), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_sort_scatter(const SortArgs& a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(sort_scatter_kernel, dim3(blocks), dim3(kSortBlock), 0, s, a);
    return hipGetLastError();
}

}  // namespace rm
