// Material sort of the shading queue: keys, bin counts and the permutation
// that groups hits by material before shading (sort.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "../device/queues.h"
#include "../device/scene.h"

namespace rm {

constexpr int kSortBlock = 256;

struct SortArgs {
    HitQueue hits;
    const PerInstance* instances;
    const uint16_t* materials;  // per face
    uint32_t* keys;             // per queue entry
    uint32_t* order;            // queue entries grouped by key
    Counters* counters;
    uint32_t iter;
    const uint32_t* iter_dev;   // captured bounce loops (iteration_of)
    uint32_t key_shift;         // granularity: materials m >> key_shift share a bin
};

// keys(iter) then scatter(iter); both read the queue length from the
// iteration's counters, so the launches need no host synchronisation.
hipError_t launch_sort_keys(const SortArgs& a, int blocks, hipStream_t s);
hipError_t launch_sort_scatter(const SortArgs& a, int blocks, hipStream_t s);

}  // namespace rm
