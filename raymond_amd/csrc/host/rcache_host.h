// Radiance-cache state behind the C ABI (include/rm.h, rm_cache_*).
#pragma once
#include "../kernels/rcache_api.h"
#include "core.h"

struct rm_cache {
    rm_device* dev = nullptr;
    rm_cache_desc desc{};
    rm::RcGrid grid{};
    size_t table_floats = 0;
    // FP32 masters, Adam moments, gradients and the FP16 copies the kernels read
    rm::DevBuf t_state;            // Adam records, 8 floats per table pair
    rm::DevBuf t_master;           // FP32 table staging for parameter reads and writes
    rm::DevBuf t_grad, t_half;
    rm::DevBuf w_master, w_m, w_v, w_grad, w_half;
    rm::DevBuf partial, loss;
    uint32_t partial_waves = 0;
    // raw-batch staging
    rm::DevBuf stage_feat, stage_target, stage_out;
    uint64_t steps = 0, samples = 0, queries = 0, records = 0;
    float last_loss = 0.0f;
    double ms_infer = 0.0, ms_train = 0.0;
};

namespace rm {

void cache_create(rm_device* dev, const rm_cache_desc* desc, rm_cache* c);
RcParams cache_params(const rm_cache* c);
// One Adam step on `batch` samples drawn from `smp` (its permutation fields
// select them), on `stream` (null: the device's stream); returns the mean
// loss only when `read_loss` (it synchronises).
float cache_step(rm_cache* c, const RcSamples& smp, uint32_t batch, bool read_loss, hipStream_t stream = nullptr);
void cache_read_params(rm_cache* c, float* mlp, size_t mlp_count, float* table, size_t table_count);
void cache_write_params(rm_cache* c, const float* mlp, size_t mlp_count, const float* table, size_t table_count);
void cache_infer_raw(rm_cache* c, const float* features, uint32_t n, float* out);
void cache_gradients_raw(rm_cache* c, const float* features, const float* targets, uint32_t n, float* grad_mlp,
                         float* grad_table, float* loss);
void cache_train_raw(rm_cache* c, const float* features, const float* targets, uint32_t n, uint32_t steps,
                     float* last_loss);

}  // namespace rm
