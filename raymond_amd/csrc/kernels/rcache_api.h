// Radiance-cache kernels (rcache.hip): inference on queued path vertices,
// training-sample targets, fused forward/backward on MFMA, the weight-gradient
// reduction and Adam.
#pragma once
#include <hip/hip_runtime.h>
#include "../device/queues.h"
#include "../device/rcache.h"

namespace rm {

// Parameters on the device.  Masters, moments and MLP gradients are FP32;
// the kernels read the FP16 copies Adam writes.  Table gradients accumulate
// as FP16 pairs (one packed atomic per grid corner) scaled by
// kRcTableGradScale, which keeps per-sample contributions of a mean loss
// over thousands of samples in FP16's normal range; a contribution is
// clamped to +-kRcTableGradClamp and Adam skips entries whose sum is not
// finite.
struct RcParams {
    RcGrid grid;
    const uint32_t* table_h;   // levels x 2^log2 entries of two FP16 features
    const uint16_t* mlp_h;     // kRcMlpParams FP16 bit patterns, [in][out] per layer
};

constexpr float kRcTableGradScale = 4096.0f;
constexpr float kRcTableGradClamp = 2048.0f;

struct RcGrads {
    uint32_t* table;           // levels x 2^log2 FP16 pairs, x kRcTableGradScale
    float* mlp_partial;        // [waves][kRcMlpParams]
    float* loss;               // sum of per-sample losses (one float)
};

// Training samples, feature-major: feat[k * stride + i] for feature k.
struct RcSamples {
    const float* feat;         // kRcFeatureFloats x stride
    const float* target;       // 3 x stride; a NaN target marks an unusable sample
    uint32_t stride;
    uint32_t count;
    uint32_t first;            // permutation: sample j of the batch is
    uint32_t perm_mul;         // record ((first + j) * perm_mul) mod count
};

struct RcInferArgs {
    RcParams p;
    RcQueryQueue q;
    const uint32_t* count_ptr; // queries of this iteration (device counter)
    uint32_t count;            // used when count_ptr is null
    uint32_t stride;           // feature-array stride (queue capacity, or count for raw batches)
    float* image;              // RGBA accumulation; null: write `out`
    float* out;                // 3 x count, feature-major (raw batches)
};

struct RcAdam {
    float* master;
    float* m;
    float* v;
    float* grad;               // MLP: FP32 gradients
    uint32_t* grad_h2;         // table: scaled FP16 gradient pairs
    uint32_t* half_out;        // FP16 copy; two values per word when pairs != 0
    uint16_t* half_mlp;        // FP16 bit patterns
    uint32_t n;
    float lr, beta1, beta2, eps;
    float bias1, bias2;        // 1 - beta^t
    float grad_scale;
};

hipError_t launch_rc_infer(const RcInferArgs& a, int blocks, hipStream_t s);
// Fused forward + backward over `batch` samples (32 per wave, 4 waves per
// workgroup): table gradients by atomic add, MLP gradients as one partial per
// wave.  `scale` multiplies every gradient (1 / batch for a mean loss).
hipError_t launch_rc_train(const RcParams& p, const RcSamples& smp, uint32_t batch, const RcGrads& g, float scale,
                           hipStream_t s);
hipError_t launch_rc_reduce(const float* partial, uint32_t waves, float* grad_mlp, hipStream_t s);
hipError_t launch_rc_adam_mlp(const RcAdam& a, hipStream_t s);
hipError_t launch_rc_adam_table(const RcAdam& a, hipStream_t s);
// Training records -> targets ((radiance at path end - r) / t).
hipError_t launch_rc_targets(const RcTrainQueue& q, const float* slot_radiance, const uint32_t* count_ptr,
                             uint32_t capacity, float* target, hipStream_t s);
hipError_t launch_rc_fill(float* p, float v, size_t n, hipStream_t s);
// FP16 copies of master parameters (after the host writes masters).
hipError_t launch_rc_halves(const float* mlp, uint16_t* mlp_h, const float* table, uint32_t* table_h,
                            size_t table_floats, hipStream_t s);
uint32_t rc_train_waves(uint32_t batch);

}  // namespace rm
