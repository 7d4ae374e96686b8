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
    // Spatially sorted records (null: the permutation above): sample j of a
    // batch of b is record order[floor((j + offset) * count / b)], a
    // systematic sample of the sorted records whose consecutive samples, and
    // so the lanes of a wave, lie close together in the scene.
    const uint32_t* order;
    float offset;              // in [0, 1), one per step
};

// Lookups: either query rows read through a spatial order (the renderer's
// path), each adding throughput x prediction to its pixel in `image` (plus
// one sample in alpha) or, without an image, writing its prediction to
// out[row]; or feature-major raw batches writing 3 x count outputs.
struct RcInferArgs {
    RcParams p;
    const float* rows;         // kRcQueryFloats per query, or null for a raw batch
    const uint32_t* order;     // rows in spatial order (null: identity)
    const float* feat;         // raw batch: kRcFeatureFloats x stride, feature-major
    uint32_t count;
    uint32_t stride;           // raw batch: feature-array stride
    float* out;                // rows: 4 floats per row (prediction, 0); raw: 3 x count
    float* image;              // rows: add throughput x prediction to the row's pixel here instead of out
};

// Spatial order of the lookups: a counting sort by the Morton code of the
// query position on a kRcSortGrid^3 grid over the scene bounds, so that the
// waves of one lookup launch gather from nearby grid cells.
constexpr uint32_t kRcSortBits = 4;                           // per axis
constexpr uint32_t kRcSortBins = 1u << (3 * kRcSortBits);     // 4096
struct RcSortArgs {
    const float* pos;          // component a of element i at pos[i * pos_elem + a * pos_comp]
    uint32_t pos_elem, pos_comp;
    uint32_t count;
    uint32_t* keys;            // per query
    uint32_t* hist;            // kRcSortBins counts, then exclusive starts
    uint32_t* cursor;          // kRcSortBins claim cursors
    uint32_t* order;           // sorted row indices
};

struct RcAdam {
    float* master;             // MLP: FP32 masters and moments
    float* m;
    float* v;
    float* state;              // table: 8 floats per pair {p0, p1, m0, m1, v0, v1, -, -}
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
// keys + histogram, exclusive scan, scatter (hist and cursor must start at 0;
// the scan leaves cursor at 0 for the scatter, and the scatter clears hist).
hipError_t launch_rc_sort_queries(const RcSortArgs& a, int blocks, hipStream_t s);
// Fused forward + backward over `batch` samples (32 per wave, two waves per
// workgroup): table gradients by atomic add, MLP gradients as one partial per
// wave.  `scale` multiplies every gradient (1 / batch for a mean loss).
hipError_t launch_rc_train(const RcParams& p, const RcSamples& smp, uint32_t batch, const RcGrads& g, float scale,
                           hipStream_t s);
hipError_t launch_rc_reduce(const float* partial, uint32_t waves, float* grad_mlp, hipStream_t s);
hipError_t launch_rc_adam_mlp(const RcAdam& a, hipStream_t s);
hipError_t launch_rc_adam_table(const RcAdam& a, hipStream_t s);
// Training records -> targets ((radiance at path end - r) / t).
hipError_t launch_rc_targets(const RcTrainQueue& q, const float* slot_radiance, uint32_t count, uint32_t capacity,
                             float* target, hipStream_t s);
hipError_t launch_rc_fill(float* p, float v, size_t n, hipStream_t s);
// FP16 copy of the MLP masters (after the host writes them).
hipError_t launch_rc_halves(const float* mlp, uint16_t* mlp_h, hipStream_t s);
// Table masters into the Adam records (and the FP16 copy) or out of them.
hipError_t launch_rc_table_master(float* state, float* master, uint32_t* table_h, size_t table_floats, bool set,
                                  hipStream_t s);
uint32_t rc_train_waves(uint32_t batch);

}  // namespace rm
