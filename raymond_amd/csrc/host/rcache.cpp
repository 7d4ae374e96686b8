// Radiance cache, host side: parameter set-up, the Adam step sequence and the
// raw-batch entry points.
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "rcache_host.h"

namespace rm {

namespace {

constexpr float kBeta1 = 0.9f, kBeta2 = 0.99f, kAdamEps = 1e-15f;

// splitmix64: parameter initialisation only
struct InitRng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uniform(float a) { return a * (2.0f * (float)((next() >> 40) * (1.0 / 16777216.0)) - 1.0f); }
};

// IEEE binary16 bit pattern to float (exact).
float half_bits_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float v;
    if (e == 0u) {
        v = std::ldexp((float)m, -24);  // zero and subnormals
    } else if (e == 31u) {
        v = m ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
    } else {
        v = std::ldexp((float)(m | 0x400u), (int)e - 25);
    }
    return sign ? -v : v;
}

// Feature-major copy of n x k row-major host rows.
std::vector<float> to_columns(const float* rows, uint32_t n, int k) {
    std::vector<float> cols((size_t)n * k);
    for (uint32_t i = 0; i < n; i++)
        for (int j = 0; j < k; j++) cols[(size_t)j * n + i] = rows[(size_t)i * k + j];
    return cols;
}

void zero_grads(rm_cache* c, hipStream_t s) {
    RM_HIP(launch_rc_fill(c->t_grad.as<float>(), 0.0f, c->table_floats / 2, s));  // FP16 pairs
    RM_HIP(launch_rc_fill(c->w_grad.as<float>(), 0.0f, kRcMlpParams, s));
}

void ensure_partial(rm_cache* c, uint32_t batch) {
    const uint32_t waves = rc_train_waves(batch);
    if (waves > c->partial_waves) {
        c->partial.alloc((size_t)waves * kRcMlpParams * sizeof(float));
        c->partial_waves = waves;
    }
}

// Forward + backward into the gradient buffers (which must be zero).
void cache_backward(rm_cache* c, const RcSamples& smp, uint32_t batch, hipStream_t s) {
    ensure_partial(c, batch);
    RcGrads g;
    g.table = c->t_grad.as<uint32_t>();
    g.mlp_partial = c->partial.as<float>();
    g.loss = c->loss.as<float>();
    RM_HIP(launch_rc_train(cache_params(c), smp, batch, g, 1.0f / (float)batch, s));
    RM_HIP(launch_rc_reduce(c->partial.as<float>(), rc_train_waves(batch), c->w_grad.as<float>(), s));
}

}  // namespace

void cache_create(rm_device* dev, const rm_cache_desc* desc, rm_cache* c) {
    c->dev = dev;
    rm_cache_desc d = desc ? *desc : rm_cache_desc{};
    if (!d.log2_table_size) d.log2_table_size = 19;
    if (!(d.learning_rate > 0.0f)) d.learning_rate = 1e-2f;
    if (!d.train_shift) d.train_shift = 4;
    // 2 steps of 8192 samples per spp: the best equal-time error of the
    // schedules measured at 1080p (4 x 16384, 2 x 16384, 4 x 8192, 4 x 4096)
    if (!d.batch_size) d.batch_size = 8192;
    if (!d.steps_per_launch) d.steps_per_launch = 2;
    RM_REQUIRE(d.log2_table_size >= 8 && d.log2_table_size <= 24, "log2_table_size must be in [8, 24]");
    RM_REQUIRE(d.train_shift <= 12, "train_shift must be <= 12");
    RM_REQUIRE(d.batch_size >= 32 && d.batch_size <= (1u << 20), "batch_size must be in [32, 2^20]");
    c->desc = d;

    // resolutions geometric from 16 to 2048 over the 16 levels
    const double b = std::exp((std::log(2048.0) - std::log(16.0)) / (kRcLevels - 1));
    const uint64_t T = 1ull << d.log2_table_size;
    c->grid.log2_table = d.log2_table_size;
    for (int l = 0; l < kRcLevels; l++) {
        const uint32_t res = (uint32_t)std::floor(16.0 * std::pow(b, l) + 0.5);
        c->grid.res[l] = res;
        const uint64_t n = res + 1ull;
        c->grid.dense[l] = n * n * n <= T ? 1u : 0u;
    }
    c->table_floats = (size_t)kRcLevels * T * 2;

    // initialisation: table U(-1e-4, 1e-4); Glorot-uniform weights over the
    // inputs in use (41 features + the constant), padding stays zero
    InitRng rng{0x5DEECE66Dull ^ ((uint64_t)d.seed << 17)};
    std::vector<float> table(c->table_floats), mlp(kRcMlpParams, 0.0f);
    for (auto& v : table) v = rng.uniform(1e-4f);
    const int used_in = kRcEncoded + kRcExtra + 1;
    const float a1 = std::sqrt(6.0f / (float)(used_in + kRcHidden));
    const float a2 = std::sqrt(6.0f / (float)(2 * kRcHidden));
    const float a3 = std::sqrt(6.0f / (float)(kRcHidden + kRcOutputs));
    for (int i = 0; i < used_in; i++)
        for (int o = 0; o < kRcHidden; o++) mlp[(size_t)i * kRcHidden + o] = rng.uniform(a1);
    for (int i = 0; i < kRcW2; i++) mlp[kRcW1 + i] = rng.uniform(a2);
    for (int i = 0; i < kRcHidden; i++)
        for (int o = 0; o < kRcOutputs; o++) mlp[kRcW1 + kRcW2 + (size_t)i * kRcW3Cols + o] = rng.uniform(a3);

    hipStream_t s = dev->stream;
    c->t_master.alloc(c->table_floats * 4);
    c->t_state.alloc(c->table_floats * 4 * 4);
    c->t_grad.alloc(c->table_floats * 2);
    c->t_half.alloc(c->table_floats * 2);
    c->w_master.alloc(kRcMlpParams * 4);
    c->w_m.alloc(kRcMlpParams * 4);
    c->w_v.alloc(kRcMlpParams * 4);
    c->w_grad.alloc(kRcMlpParams * 4);
    c->w_half.alloc(kRcMlpParams * 2);
    c->loss.alloc(4);
    RM_HIP(launch_rc_fill(c->t_state.as<float>(), 0.0f, c->table_floats * 4, s));
    RM_HIP(launch_rc_fill(c->w_m.as<float>(), 0.0f, kRcMlpParams, s));
    RM_HIP(launch_rc_fill(c->w_v.as<float>(), 0.0f, kRcMlpParams, s));
    zero_grads(c, s);
    cache_write_params(c, mlp.data(), mlp.size(), table.data(), table.size());
}

RcParams cache_params(const rm_cache* c) {
    RcParams p;
    p.grid = c->grid;
    p.table_h = c->t_half.as<uint32_t>();
    p.mlp_h = c->w_half.as<uint16_t>();
    return p;
}

float cache_step(rm_cache* c, const RcSamples& smp, uint32_t batch, bool read_loss, hipStream_t stream) {
    hipStream_t s = stream ? stream : c->dev->stream;
    RM_HIP(launch_rc_fill(c->loss.as<float>(), 0.0f, 1, s));
    cache_backward(c, smp, batch, s);
    c->steps++;
    c->samples += batch;
    RcAdam a{};
    a.lr = c->desc.learning_rate;
    a.beta1 = kBeta1;
    a.beta2 = kBeta2;
    a.eps = kAdamEps;
    a.bias1 = (float)(1.0 - std::pow((double)kBeta1, (double)c->steps));
    a.bias2 = (float)(1.0 - std::pow((double)kBeta2, (double)c->steps));
    a.grad_scale = 1.0f;
    a.master = c->w_master.as<float>();
    a.m = c->w_m.as<float>();
    a.v = c->w_v.as<float>();
    a.grad = c->w_grad.as<float>();
    a.half_mlp = c->w_half.as<uint16_t>();
    a.n = kRcMlpParams;
    RM_HIP(launch_rc_adam_mlp(a, s));
    a.master = a.m = a.v = nullptr;
    a.state = c->t_state.as<float>();
    a.grad = nullptr;
    a.grad_h2 = c->t_grad.as<uint32_t>();
    a.half_out = c->t_half.as<uint32_t>();
    a.n = (uint32_t)c->table_floats;
    RM_HIP(launch_rc_adam_table(a, s));
    if (!read_loss) return c->last_loss;
    float l = 0.0f;
    RM_HIP(hipMemcpyAsync(&l, c->loss.ptr, 4, hipMemcpyDeviceToHost, s));
    RM_HIP(hipStreamSynchronize(s));
    c->last_loss = l;
    return l;
}

void cache_read_params(rm_cache* c, float* mlp, size_t mlp_count, float* table, size_t table_count) {
    hipStream_t s = c->dev->stream;
    if (mlp) {
        RM_REQUIRE(mlp_count >= (size_t)kRcMlpParams, "mlp buffer too small");
        RM_HIP(hipMemcpyAsync(mlp, c->w_master.ptr, kRcMlpParams * 4, hipMemcpyDeviceToHost, s));
    }
    if (table) {
        RM_REQUIRE(table_count >= c->table_floats, "table buffer too small");
        RM_HIP(launch_rc_table_master(c->t_state.as<float>(), c->t_master.as<float>(), c->t_half.as<uint32_t>(),
                                      c->table_floats, false, s));
        RM_HIP(hipMemcpyAsync(table, c->t_master.ptr, c->table_floats * 4, hipMemcpyDeviceToHost, s));
    }
    RM_HIP(hipStreamSynchronize(s));
}

void cache_write_params(rm_cache* c, const float* mlp, size_t mlp_count, const float* table, size_t table_count) {
    hipStream_t s = c->dev->stream;
    if (mlp) {
        RM_REQUIRE(mlp_count >= (size_t)kRcMlpParams, "mlp buffer too small");
        RM_HIP(hipMemcpyAsync(c->w_master.ptr, mlp, kRcMlpParams * 4, hipMemcpyHostToDevice, s));
    }
    if (table) {
        RM_REQUIRE(table_count >= c->table_floats, "table buffer too small");
        RM_HIP(hipMemcpyAsync(c->t_master.ptr, table, c->table_floats * 4, hipMemcpyHostToDevice, s));
        RM_HIP(launch_rc_table_master(c->t_state.as<float>(), c->t_master.as<float>(), c->t_half.as<uint32_t>(),
                                      c->table_floats, true, s));
    }
    RM_HIP(launch_rc_halves(c->w_master.as<float>(), c->w_half.as<uint16_t>(), s));
    RM_HIP(hipStreamSynchronize(s));
}

void cache_infer_raw(rm_cache* c, const float* features, uint32_t n, float* out) {
    if (n == 0) return;
    hipStream_t s = c->dev->stream;
    const std::vector<float> cols = to_columns(features, n, kRcFeatureFloats);
    c->stage_feat.upload(cols.data(), cols.size() * 4, s);
    c->stage_out.alloc((size_t)n * 3 * 4);
    RcInferArgs a{};
    a.p = cache_params(c);
    a.feat = c->stage_feat.as<float>();
    a.count = n;
    a.stride = n;
    a.out = c->stage_out.as<float>();
    const uint32_t blocks = std::min<uint32_t>((n + 127) / 128, 4096);
    RM_HIP(launch_rc_infer(a, (int)blocks, s));
    std::vector<float> res((size_t)n * 3);
    RM_HIP(hipMemcpyAsync(res.data(), c->stage_out.ptr, res.size() * 4, hipMemcpyDeviceToHost, s));
    RM_HIP(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) out[(size_t)i * 3 + k] = res[(size_t)k * n + i];
}

static RcSamples stage_samples(rm_cache* c, const float* features, const float* targets, uint32_t n) {
    hipStream_t s = c->dev->stream;
    const std::vector<float> f = to_columns(features, n, kRcFeatureFloats);
    const std::vector<float> t = to_columns(targets, n, 3);
    c->stage_feat.upload(f.data(), f.size() * 4, s);
    c->stage_target.upload(t.data(), t.size() * 4, s);
    RcSamples smp{};
    smp.feat = c->stage_feat.as<float>();
    smp.target = c->stage_target.as<float>();
    smp.stride = n;
    smp.count = n;
    smp.first = 0;
    smp.perm_mul = 1;
    return smp;
}

void cache_gradients_raw(rm_cache* c, const float* features, const float* targets, uint32_t n, float* grad_mlp,
                         float* grad_table, float* loss) {
    RM_REQUIRE(n > 0, "empty batch");
    hipStream_t s = c->dev->stream;
    const RcSamples smp = stage_samples(c, features, targets, n);
    RM_HIP(launch_rc_fill(c->loss.as<float>(), 0.0f, 1, s));
    cache_backward(c, smp, n, s);
    if (grad_mlp) RM_HIP(hipMemcpyAsync(grad_mlp, c->w_grad.ptr, kRcMlpParams * 4, hipMemcpyDeviceToHost, s));
    std::vector<uint16_t> bits(grad_table ? c->table_floats : 0);
    if (grad_table) RM_HIP(hipMemcpyAsync(bits.data(), c->t_grad.ptr, c->table_floats * 2, hipMemcpyDeviceToHost, s));
    if (loss) RM_HIP(hipMemcpyAsync(loss, c->loss.ptr, 4, hipMemcpyDeviceToHost, s));
    zero_grads(c, s);
    RM_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i < bits.size(); i++) grad_table[i] = half_bits_to_float(bits[i]) / kRcTableGradScale;
}

void cache_train_raw(rm_cache* c, const float* features, const float* targets, uint32_t n, uint32_t steps,
                     float* last_loss) {
    RM_REQUIRE(n > 0, "empty batch");
    const RcSamples smp = stage_samples(c, features, targets, n);
    float l = 0.0f;
    for (uint32_t k = 0; k < steps; k++) l = cache_step(c, smp, n, k + 1 == steps);
    if (last_loss) *last_loss = l;
}

}  // namespace rm
