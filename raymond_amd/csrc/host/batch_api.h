// Batch state behind the C ABI (include/rm.h, rm_batch_*).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "core.h"

struct rm_batch {
    rm_device* dev = nullptr;
    std::vector<rm_job_desc> desc;
    rm_scene merged;                      // the jobs' scenes as one device scene
    std::unique_ptr<rm_renderer> renderer;
    size_t renderer_capacity = 0;
    std::vector<rm::JobInfo> jobs;
    std::vector<uint64_t> job_paths;      // width x height x samples per job
    std::vector<rm::BatchSlice> slices;   // sample-major, firsts filled per bounce loop
    std::vector<rm::BatchSlice> slices_sjf;  // shortest job first (fewest paths), samples in order
    std::vector<double> job_ms;           // completion time of each job in the last render
    rm::DevBuf d_jobs, d_lights, d_done, d_slices;
    uint32_t pixels = 0;
    uint32_t max_depth = 1;
};

namespace rm {

void batch_create(rm_device* dev, const rm_job_desc* jobs, uint32_t n, rm_batch* b);
// Renders the first `max_slices` slices (all when 0) with `cfg`.
void batch_render(rm_batch* b, const rm_batch_config* cfg, rm_batch_stats* stats, size_t max_slices = 0);
void batch_read_image(rm_batch* b, uint32_t job, float* rgba, size_t count);
void batch_autotune(rm_batch* b, rm_batch_config* best, rm_batch_tune_entry* entries, uint32_t capacity,
                    uint32_t* count, int* cache_hit);

}  // namespace rm
