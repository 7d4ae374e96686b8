// Radiance-cache kernels for gfx950.
//
// One wave evaluates 32 samples.  The two halves of the wave split the
// encoding: lane (n, hi) with n = lane % 32, hi = lane / 32 interpolates the
// 8 hash-grid levels {4c + 2hi, 4c + 2hi + 1 : c = 0..3} of sample n.  The
// MLP runs transposed, H^T = W^T X^T, on v_mfma_f32_32x32x8_f16: with the
// batch as the N dimension, the accumulator of one layer (lane (n, hi), value
// a = feature 8(a/4) + 4hi + a%4 of sample n) is exactly the B operand the
// next layer needs (lane (n, hi), k = 8c + 4hi + v), so activations go from
// layer to layer in registers.  The same holds for the encoding: the B
// operand chunks 0..3 are the features of the levels the lane interpolated,
// and dL/dx comes back in the lanes that scatter it into the table.
// Weight gradients sum over the batch, which is the K dimension of
// dW = A B^T; activations go through LDS ([feature][sample]) for that
// transpose.
#include "rcache_api.h"

namespace rm {

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float acc16_t __attribute__((ext_vector_type(16)));

constexpr int kInferWaves = 4;
// Waves per training workgroup.  A step of 8192 samples is 256 waves; the
// kernel waits on its table-gradient atomics, so more, smaller workgroups
// keep more of them in flight: per step 119.8 us at 8 waves, 95.7 at 4,
// 84.4 at 2, 84.0 at 1 (2 has the lower LDS cost per wave of weights).
#ifndef RM_TRAIN_WAVES
#define RM_TRAIN_WAVES 2
#endif
constexpr int kTrainWaves = RM_TRAIN_WAVES;

// LDS row strides (halves), padded against bank conflicts; multiples of 4 so
// every 4-half operand read is one aligned 8-byte load.
constexpr int kS1 = 68;   // W1  [48 in][64 out]
constexpr int kS1T = 52;  // W1T [64 out][48 in]
constexpr int kS2 = 68;   // W2, W2T [64][64]
constexpr int kS3 = 36;   // W3  [64 in][32 out]
constexpr int kS3T = 68;  // W3T [32 out][64 in]
constexpr int kSA = 36;   // activations [feature][32 samples]

constexpr int kOffW2 = kRcW1;
constexpr int kOffW3 = kRcW1 + kRcW2;

struct RcLdsFwd {
    _Float16 w1t[64 * kS1T];
    _Float16 w2t[64 * kS2];
    _Float16 w3t[32 * kS3T];
};
struct RcLdsBwd {
    _Float16 w1[48 * kS1];
    _Float16 w2[64 * kS2];
    _Float16 w3[64 * kS3];
};
// Per-wave transpose buffers of the weight gradients, reused layer by layer:
// (X, dZ1) for dW1, then (H1, dZ2) for dW2, then (H2, dY) for dW3.
struct RcLdsWave {
    _Float16 a[64 * kSA];
    _Float16 b[64 * kSA];
};

__device__ inline acc16_t mfma(half4_t a, half4_t b, acc16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, c, 0, 0, 0);
}
__device__ inline half4_t lds4(const _Float16* base, int stride, int row, int col) {
    return *reinterpret_cast<const half4_t*>(base + row * stride + col);
}
__device__ inline acc16_t zero16() {
    acc16_t z;
#pragma unroll
    for (int a = 0; a < 16; a++) z[a] = 0.0f;
    return z;
}
__device__ inline half4_t relu4(const acc16_t& z, int j) {
    half4_t h;
#pragma unroll
    for (int v = 0; v < 4; v++) h[v] = (_Float16)max_(z[4 * j + v], 0.0f);
    return h;
}
__device__ inline float half_lo(uint32_t w) {
    _Float16 h;
    const uint16_t b = (uint16_t)(w & 0xffffu);
    __builtin_memcpy(&h, &b, 2);
    return (float)h;
}
__device__ inline float half_hi(uint32_t w) {
    _Float16 h;
    const uint16_t b = (uint16_t)(w >> 16);
    __builtin_memcpy(&h, &b, 2);
    return (float)h;
}

// Calls f(i, w[i]) for the `count` FP16 values from w + offset, read 8 at a
// time (one 16-byte load per lane; offsets and counts are multiples of 8).
template <class F>
__device__ inline void for_each_half8(const uint16_t* bits, int offset, int count, F f) {
    const uint4* src = reinterpret_cast<const uint4*>(bits + offset);
    for (int j = threadIdx.x; j < count / 8; j += (int)blockDim.x) {
        const uint4 v = src[j];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const uint16_t b = (uint16_t)(w[e >> 1] >> (16 * (e & 1)));
            _Float16 h;
            __builtin_memcpy(&h, &b, 2);
            f(8 * j + e, h);
        }
    }
}

__device__ void load_fwd_weights(RcLdsFwd& W, const uint16_t* bits) {
    for_each_half8(bits, 0, kRcW1, [&](int i, _Float16 h) { W.w1t[(i % kRcHidden) * kS1T + i / kRcHidden] = h; });
    for_each_half8(bits, kOffW2, kRcW2, [&](int i, _Float16 h) { W.w2t[(i % kRcHidden) * kS2 + i / kRcHidden] = h; });
    for_each_half8(bits, kOffW3, kRcW3, [&](int i, _Float16 h) { W.w3t[(i % kRcW3Cols) * kS3T + i / kRcW3Cols] = h; });
}
__device__ void load_bwd_weights(RcLdsBwd& W, const uint16_t* bits) {
    for_each_half8(bits, 0, kRcW1, [&](int i, _Float16 h) { W.w1[(i / kRcHidden) * kS1 + i % kRcHidden] = h; });
    for_each_half8(bits, kOffW2, kRcW2, [&](int i, _Float16 h) { W.w2[(i / kRcHidden) * kS2 + i % kRcHidden] = h; });
    for_each_half8(bits, kOffW3, kRcW3, [&](int i, _Float16 h) { W.w3[(i / kRcW3Cols) * kS3 + i % kRcW3Cols] = h; });
}

// Corner c (bits x, y, z) of the cell around p at `level`: table index and
// trilinear weight.
struct RcCell {
    uint32_t ix, iy, iz;
    float fx, fy, fz;
};
__device__ inline RcCell rc_locate(const RcGrid& g, int level, vec3 p) {
    RcCell c;
    const uint32_t res = g.res[level];
    rc_cell(p.x, res, c.ix, c.fx);
    rc_cell(p.y, res, c.iy, c.fy);
    rc_cell(p.z, res, c.iz, c.fz);
    return c;
}
__device__ inline float rc_corner(const RcGrid& g, int level, const RcCell& c, int k, uint32_t& idx) {
    const uint32_t bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
    idx = rc_index(g, level, c.ix + bx, c.iy + by, c.iz + bz);
    const float wx = bx ? c.fx : 1.0f - c.fx;
    const float wy = by ? c.fy : 1.0f - c.fy;
    const float wz = bz ? c.fz : 1.0f - c.fz;
    return (wx * wy) * wz;
}

// B operand chunks 0..5 of sample n for lane half hi: the encoding of the
// lane's 8 levels, then the vertex features and the constant input.  The
// loop over level pairs stays rolled (16 gathers in flight per lane instead
// of 64 keeps the register count, and with it the occupancy, up); its result
// is placed by value selects, not by indexing a register array.  The
// per-lane branch on a level's dense flag makes the compiler issue the
// corner gathers two at a time; a branch-free form that issued each level
// pair's 16 gathers together measured slower on the cache comparison
// (equal-time ratio 0.590-0.595 against 0.575-0.581), its per-level
// parameter loads and a register spill costing more than the overlap won.
__device__ inline void rc_inputs(const RcParams& p, const float* f, int hi, half4_t x[6]) {
    const vec3 pos(f[0], f[1], f[2]);
    half4_t x0 = {}, x1 = {}, x2 = {}, x3 = {};
#pragma unroll 1
    for (int c = 0; c < 4; c++) {
        half4_t v;
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int level = 4 * c + 2 * hi + j;
            const RcCell cell = rc_locate(p.grid, level, pos);
            const uint32_t* tab = p.table_h + ((size_t)level << p.grid.log2_table);
            float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                uint32_t idx;
                const float w = rc_corner(p.grid, level, cell, k, idx);
                const uint32_t word = tab[idx];
                f0 += w * half_lo(word);
                f1 += w * half_hi(word);
            }
            v[2 * j] = (_Float16)f0;
            v[2 * j + 1] = (_Float16)f1;
        }
        x0 = c == 0 ? v : x0;
        x1 = c == 1 ? v : x1;
        x2 = c == 2 ? v : x2;
        x3 = c == 3 ? v : x3;
    }
    x[0] = x0;
    x[1] = x1;
    x[2] = x2;
    x[3] = x3;
    // features 32 + 4hi + v (chunk 4) and 40 + 4hi + v (chunk 5)
#pragma unroll
    for (int v = 0; v < 4; v++) x[4][v] = (_Float16)f[3 + 4 * hi + v];
    x[5][0] = (_Float16)(hi ? 0.0f : f[3 + 8]);
    x[5][1] = (_Float16)(hi ? 0.0f : 1.0f);
    x[5][2] = (_Float16)0.0f;
    x[5][3] = (_Float16)0.0f;
}

struct RcAct {
    half4_t h1[8], h2[8];
    acc16_t y;
};

__device__ inline void rc_forward(const RcLdsFwd& W, const half4_t x[6], int n, int hi, RcAct& act) {
#pragma unroll
    for (int t = 0; t < 2; t++) {
        acc16_t z = zero16();
#pragma unroll
        for (int c = 0; c < 6; c++) z = mfma(lds4(W.w1t, kS1T, 32 * t + n, 8 * c + 4 * hi), x[c], z);
#pragma unroll
        for (int j = 0; j < 4; j++) act.h1[4 * t + j] = relu4(z, j);
    }
#pragma unroll
    for (int t = 0; t < 2; t++) {
        acc16_t z = zero16();
#pragma unroll
        for (int c = 0; c < 8; c++) z = mfma(lds4(W.w2t, kS2, 32 * t + n, 8 * c + 4 * hi), act.h1[c], z);
#pragma unroll
        for (int j = 0; j < 4; j++) act.h2[4 * t + j] = relu4(z, j);
    }
    acc16_t y = zero16();
#pragma unroll
    for (int c = 0; c < 8; c++) y = mfma(lds4(W.w3t, kS3T, n, 8 * c + 4 * hi), act.h2[c], y);
    act.y = y;
}

// Three waves per SIMD is the most the register budget allows without spills
// (167 VGPRs).  Forcing 4 (128 VGPRs, 112 bytes of spills per lane) or 5
// (96, 240 bytes) lost with the L2-resident 2^15 table: equal-time cache
// ratio 0.601-0.611 and 0.624 against 0.589-0.591.
// Query rows are read through the spatial order; when the grid is a multiple
// of 8 workgroups, the tiles are cut into 8 contiguous ranges, one per
// workgroup residue mod 8 (consecutive workgroups go to different XCDs), so
// that each XCD's L2 serves one region of the scene.
__global__ __launch_bounds__(64 * kInferWaves, 3) void rc_infer_kernel(RcInferArgs a) {
    __shared__ RcLdsFwd W;
    load_fwd_weights(W, a.p.mlp_h);
    __syncthreads();
    typedef float row4 __attribute__((ext_vector_type(4)));
    const uint32_t count = a.count;
    const int lane = threadIdx.x & 63, n = lane & 31, hi = lane >> 5;
    const uint32_t tiles = (count + 31u) / 32u;
    const uint32_t wave = threadIdx.x / 64u;
    uint32_t t1 = tiles, first = blockIdx.x * kInferWaves + wave, step = gridDim.x * kInferWaves;
    if (a.rows && (gridDim.x & 7u) == 0u) {
        const uint32_t part = blockIdx.x & 7u, per = (tiles + 7u) / 8u;
        const uint32_t t0 = part * per;
        t1 = min(tiles, t0 + per);
        first = t0 + (blockIdx.x >> 3) * kInferWaves + wave;
        step = (gridDim.x >> 3) * kInferWaves;
    }
    for (uint32_t tile = first; tile < t1; tile += step) {
        const uint32_t i = tile * 32u + (uint32_t)n;
        const bool valid = i < count;
        float f[kRcFeatureFloats];
        uint32_t row = i;
        if (a.rows) {
            row = valid ? (a.order ? a.order[i] : i) : 0u;
            const row4* r = reinterpret_cast<const row4*>(a.rows + (size_t)row * kRcQueryFloats);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const row4 v = valid ? r[1 + c] : row4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int k = 0; k < 4; k++) f[4 * c + k] = v[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < kRcFeatureFloats; k++) f[k] = valid ? a.feat[(size_t)k * a.stride + i] : 0.0f;
        }
        half4_t x[6];
        rc_inputs(a.p, f, hi, x);
        RcAct act;
        rc_forward(W, x, n, hi, act);
        if (hi == 0 && valid) {
            if (a.rows && a.image) {
                // the lookup's contribution straight into its pixel: throughput
                // x prediction, and one sample in alpha
                const row4 r0 = reinterpret_cast<const row4*>(a.rows + (size_t)row * kRcQueryFloats)[0];
                float* px = a.image + 4u * (size_t)__builtin_bit_cast(uint32_t, r0[0]);
                atomicAdd(px + 0, r0[1] * max_(act.y[0], 0.0f));
                atomicAdd(px + 1, r0[2] * max_(act.y[1], 0.0f));
                atomicAdd(px + 2, r0[3] * max_(act.y[2], 0.0f));
                atomicAdd(px + 3, 1.0f);
            } else if (a.rows) {
                row4 v;
                v[0] = max_(act.y[0], 0.0f);
                v[1] = max_(act.y[1], 0.0f);
                v[2] = max_(act.y[2], 0.0f);
                v[3] = 0.0f;
                reinterpret_cast<row4*>(a.out)[row] = v;
            } else {
                a.out[i] = act.y[0];
                a.out[(size_t)count + i] = act.y[1];
                a.out[2 * (size_t)count + i] = act.y[2];
            }
        }
    }
}

// Morton code of a position in [0, 1]^3 on the sort grid.
__device__ inline uint32_t rc_sort_key(const RcSortArgs& s, uint32_t i) {
    constexpr float cells = (float)(1u << kRcSortBits);
    uint32_t key = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float v = s.pos[(size_t)i * s.pos_elem + (size_t)a * s.pos_comp] * cells;
        uint32_t c = v > 0.0f ? (uint32_t)v : 0u;
        c = c < (1u << kRcSortBits) ? c : (1u << kRcSortBits) - 1u;
#pragma unroll
        for (uint32_t b = 0; b < kRcSortBits; b++) key |= ((c >> b) & 1u) << (3u * b + (uint32_t)a);
    }
    return key;
}

constexpr int kRcSortBlock = 256;

__global__ __launch_bounds__(kRcSortBlock) void rc_qkeys_kernel(RcSortArgs a) {
    __shared__ uint32_t hist[kRcSortBins];
    for (uint32_t b = threadIdx.x; b < kRcSortBins; b += kRcSortBlock) hist[b] = 0u;
    __syncthreads();
    for (uint32_t i = blockIdx.x * kRcSortBlock + threadIdx.x; i < a.count; i += gridDim.x * kRcSortBlock) {
        const uint32_t k = rc_sort_key(a, i);
        a.keys[i] = k;
        atomicAdd(&hist[k], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kRcSortBins; b += kRcSortBlock)
        if (hist[b]) atomicAdd(&a.hist[b], hist[b]);
}

// One workgroup: exclusive prefix of the bin counts, in place.
__global__ __launch_bounds__(1024) void rc_qscan_kernel(RcSortArgs a) {
    constexpr uint32_t per = kRcSortBins / 1024u;
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    uint32_t v[per], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < per; k++) {
        v[k] = a.hist[t * per + k];
        sum += v[k];
    }
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        const uint32_t add = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
#pragma unroll
    for (uint32_t k = 0; k < per; k++) {
        a.hist[t * per + k] = run;
        run += v[k];
    }
}

__global__ __launch_bounds__(kRcSortBlock) void rc_qscatter_kernel(RcSortArgs a) {
    __shared__ uint32_t cnt[kRcSortBins];
    __shared__ uint32_t base[kRcSortBins];
    for (uint32_t b = threadIdx.x; b < kRcSortBins; b += kRcSortBlock) cnt[b] = 0u;
    __syncthreads();
    const uint32_t stride = gridDim.x * kRcSortBlock;
    for (uint32_t i = blockIdx.x * kRcSortBlock + threadIdx.x; i < a.count; i += stride) atomicAdd(&cnt[a.keys[i]], 1u);
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < kRcSortBins; b += kRcSortBlock) {
        const uint32_t c = cnt[b];
        base[b] = c ? a.hist[b] + atomicAdd(&a.cursor[b], c) : 0u;
        cnt[b] = 0u;
    }
    __syncthreads();
    for (uint32_t i = blockIdx.x * kRcSortBlock + threadIdx.x; i < a.count; i += stride) {
        const uint32_t k = a.keys[i];
        a.order[base[k] + atomicAdd(&cnt[k], 1u)] = i;
    }
}

// Clears the starts and cursors after the scatter, for the next sort.
__global__ void rc_qclear_kernel(RcSortArgs a) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < kRcSortBins; b += gridDim.x * blockDim.x) {
        a.hist[b] = 0u;
        a.cursor[b] = 0u;
    }
}

// Clamp of a scaled table-gradient contribution; NaN passes through, so the
// pair it lands in is dropped by Adam.
__device__ inline float clamp_grad(float x) {
    return x > kRcTableGradClamp ? kRcTableGradClamp : (x < -kRcTableGradClamp ? -kRcTableGradClamp : x);
}

// Orders a wave's LDS writes before its following reads (and the reads of
// one layer before the next layer's writes to the same buffers).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Writes B-operand chunk `c` of lane (n, hi) into [feature][sample] LDS.
__device__ inline void store_chunk(_Float16* base, int c, int n, int hi, half4_t v) {
#pragma unroll
    for (int k = 0; k < 4; k++) base[(8 * c + 4 * hi + k) * kSA + n] = v[k];
}

// dW[rows][cols] += A[rows][samples] B[cols][samples]^T for one 32x32 tile.
__device__ inline acc16_t outer_tile(const _Float16* A, int ta, const _Float16* B, int tb, int n, int hi) {
    acc16_t acc = zero16();
#pragma unroll
    for (int c = 0; c < 4; c++)
        acc = mfma(lds4(A, kSA, 32 * ta + n, 8 * c + 4 * hi), lds4(B, kSA, 32 * tb + n, 8 * c + 4 * hi), acc);
    return acc;
}
// Stores a tile of dW (rows = inputs, cols = outputs) into the wave's partial.
__device__ inline void store_grad_tile(float* dst, int ld, int rows, int cols_valid, const acc16_t& acc, int ta,
                                       int tb, int n, int hi, float scale) {
    const int col = 32 * tb + n;
#pragma unroll
    for (int a = 0; a < 16; a++) {
        const int row = 32 * ta + 8 * (a / 4) + 4 * hi + (a % 4);
        if (row < rows) dst[row * ld + col] = (n < cols_valid) ? acc[a] * scale : 0.0f;
    }
}

__global__ __launch_bounds__(64 * kTrainWaves) void rc_train_kernel(RcParams p, RcSamples smp, uint32_t batch, RcGrads g,
                                                             float scale) {
    __shared__ RcLdsFwd WF;
    __shared__ RcLdsBwd WB;
    __shared__ RcLdsWave WV[kTrainWaves];
    load_fwd_weights(WF, p.mlp_h);
    load_bwd_weights(WB, p.mlp_h);
    const int wave = threadIdx.x / 64;
    RcLdsWave& L = WV[wave];
    const int lane = threadIdx.x & 63, n = lane & 31, hi = lane >> 5;
    // rows the dW1 tiles read but no lane writes: input features 48..63
    for (int k = lane; k < 16 * kSA; k += 64) L.a[48 * kSA + k] = (_Float16)0.0f;
    __syncthreads();

    const uint32_t gwave = blockIdx.x * kTrainWaves + wave;
    const uint32_t j = gwave * 32u + (uint32_t)n;
    const bool valid = j < batch && smp.count > 0;
    uint32_t rec = 0;
    if (valid) {
        if (smp.order) {
            const uint32_t q = (uint32_t)(((double)j + (double)smp.offset) * (double)smp.count / (double)batch);
            rec = smp.order[q < smp.count ? q : smp.count - 1u];
        } else {
            rec = (uint32_t)((((uint64_t)smp.first + j) * smp.perm_mul) % smp.count);
        }
    }
    float f[kRcFeatureFloats];
#pragma unroll
    for (int k = 0; k < kRcFeatureFloats; k++) f[k] = valid ? smp.feat[(size_t)k * smp.stride + rec] : 0.0f;
    half4_t x[6];
    rc_inputs(p, f, hi, x);
    RcAct act;
    rc_forward(WF, x, n, hi, act);

    // relative L2 loss and its gradient (channels 0..2 live in the lanes hi = 0)
    float dy[3] = {0.0f, 0.0f, 0.0f};
    if (hi == 0 && valid) {
        float loss = 0.0f;
        bool ok = true;
        float t[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            t[c] = smp.target[(size_t)c * smp.stride + rec];
            ok = ok && isfinite_(t[c]);
        }
        if (ok) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float y = act.y[c];
                const float den = y * y + 0.01f;
                const float d = y - t[c];
                loss += (d * d) / den;
                dy[c] = (2.0f * d) / den;
            }
            atomicAdd(g.loss, loss * scale);
        }
    }
    half4_t dyb;
    dyb[0] = (_Float16)dy[0];
    dyb[1] = (_Float16)dy[1];
    dyb[2] = (_Float16)dy[2];
    dyb[3] = (_Float16)0.0f;

    // dH2 = W3 dY, dZ2 = dH2 * [H2 > 0]
    half4_t dz2[8];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const acc16_t d = mfma(lds4(WB.w3, kS3, 32 * t + n, 4 * hi), dyb, zero16());
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                dz2[4 * t + jj][v] = act.h2[4 * t + jj][v] > (_Float16)0.0f ? (_Float16)d[4 * jj + v] : (_Float16)0.0f;
    }
    // dH1 = W2 dZ2, dZ1 = dH1 * [H1 > 0]
    half4_t dz1[8];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        acc16_t d = zero16();
#pragma unroll
        for (int c = 0; c < 8; c++) d = mfma(lds4(WB.w2, kS2, 32 * t + n, 8 * c + 4 * hi), dz2[c], d);
#pragma unroll
        for (int jj = 0; jj < 4; jj++)
#pragma unroll
            for (int v = 0; v < 4; v++)
                dz1[4 * t + jj][v] = act.h1[4 * t + jj][v] > (_Float16)0.0f ? (_Float16)d[4 * jj + v] : (_Float16)0.0f;
    }
    // dX (encoding features only) = W1[0:32] dZ1, scattered into the table
    {
        acc16_t d = zero16();
#pragma unroll
        for (int c = 0; c < 8; c++) d = mfma(lds4(WB.w1, kS1, n, 8 * c + 4 * hi), dz1[c], d);
        if (valid) {
            const vec3 pos(f[0], f[1], f[2]);
#pragma unroll
            for (int c = 0; c < 4; c++) {
#pragma unroll
                for (int jj = 0; jj < 2; jj++) {
                    const int level = 4 * c + 2 * hi + jj;
                    const float gs = scale * kRcTableGradScale;
                    const float g0 = d[4 * c + 2 * jj] * gs, g1 = d[4 * c + 2 * jj + 1] * gs;
                    if (g0 == 0.0f && g1 == 0.0f) continue;
                    const RcCell cell = rc_locate(p.grid, level, pos);
                    uint32_t* gt = g.table + ((size_t)level << p.grid.log2_table);
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        uint32_t idx;
                        const float w = rc_corner(p.grid, level, cell, k, idx);
                        half2_t v;
                        v[0] = (_Float16)clamp_grad(w * g0);
                        v[1] = (_Float16)clamp_grad(w * g1);
                        typedef __attribute__((address_space(1))) half2_t global_half2;
                        __builtin_amdgcn_global_atomic_fadd_v2f16((global_half2*)(gt + idx), v);
                    }
                }
            }
        }
    }

    // weight gradients: activations to [feature][sample] LDS, then dW = A B^T,
    // one layer at a time through the wave's two buffers
    float* part = g.mlp_partial + (size_t)gwave * kRcMlpParams;
#pragma unroll
    for (int c = 0; c < 6; c++) store_chunk(L.a, c, n, hi, x[c]);
#pragma unroll
    for (int c = 0; c < 8; c++) store_chunk(L.b, c, n, hi, dz1[c]);
    wave_sync();
#pragma unroll
    for (int ta = 0; ta < 2; ta++)
#pragma unroll
        for (int tb = 0; tb < 2; tb++)
            store_grad_tile(part, kRcHidden, kRcInputs, 32, outer_tile(L.a, ta, L.b, tb, n, hi), ta, tb, n, hi, scale);
    wave_sync();
#pragma unroll
    for (int c = 0; c < 8; c++) {
        store_chunk(L.a, c, n, hi, act.h1[c]);
        store_chunk(L.b, c, n, hi, dz2[c]);
    }
    wave_sync();
#pragma unroll
    for (int ta = 0; ta < 2; ta++)
#pragma unroll
        for (int tb = 0; tb < 2; tb++)
            store_grad_tile(part + kOffW2, kRcHidden, kRcHidden, 32, outer_tile(L.a, ta, L.b, tb, n, hi), ta, tb, n,
                            hi, scale);
    wave_sync();
    // dY occupies rows 0..7 of b; rows 8..31 hold finite values of dZ2 that
    // only reach the output columns >= 3, which are discarded
#pragma unroll
    for (int c = 0; c < 8; c++) store_chunk(L.a, c, n, hi, act.h2[c]);
    store_chunk(L.b, 0, n, hi, dyb);
    wave_sync();
#pragma unroll
    for (int ta = 0; ta < 2; ta++)
        store_grad_tile(part + kOffW3, kRcW3Cols, kRcHidden, kRcOutputs, outer_tile(L.a, ta, L.b, 0, n, hi), ta, 0, n,
                        hi, scale);
}

// Sums the per-wave partials: blockIdx.y takes a slice of kReduceSlice waves
// and adds its sum to the gradient atomically.
constexpr uint32_t kReduceSlice = 16;
__global__ void rc_reduce_kernel(const float* partial, uint32_t waves, float* grad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)kRcMlpParams) return;
    const uint32_t w0 = blockIdx.y * kReduceSlice;
    const uint32_t w1 = w0 + kReduceSlice < waves ? w0 + kReduceSlice : waves;
    float s = 0.0f;
    for (uint32_t w = w0; w < w1; w++) s += partial[(size_t)w * kRcMlpParams + i];
    atomicAdd(grad + i, s);
}

__device__ inline float adam_step(const RcAdam& a, size_t i, float g) {
    const float m = a.beta1 * a.m[i] + (1.0f - a.beta1) * g;
    const float v = a.beta2 * a.v[i] + (1.0f - a.beta2) * (g * g);
    a.m[i] = m;
    a.v[i] = v;
    const float mh = m / a.bias1;
    const float vh = v / a.bias2;
    const float p = a.master[i] - a.lr * mh / (sqrt_(vh) + a.eps);
    a.master[i] = p;
    return p;
}

__global__ void rc_adam_mlp_kernel(RcAdam a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float g = a.grad[i] * a.grad_scale;
    a.grad[i] = 0.0f;
    const _Float16 h = (_Float16)adam_step(a, i, g);
    uint16_t b;
    __builtin_memcpy(&b, &h, 2);
    a.half_mlp[i] = b;
}

// Table entries: two features per 32-bit FP16 pair, and per pair one
// 32-byte Adam record {p0, p1, m0, m1, v0, v1, -, -} (one sector per touched
// entry).  Entries no sample of the step touched keep their parameters and
// moments (lazy Adam: the moments of a sparse row decay only when the row is
// used), so a step costs a read of the gradient plus the touched entries.  A
// pair whose scaled sum overflowed is dropped for this step.
__global__ void rc_adam_table_kernel(RcAdam a) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const uint32_t stride = gridDim.x * blockDim.x;
    const float unscale = a.grad_scale / kRcTableGradScale;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < a.n / 2u; e += stride) {
        const uint32_t word = a.grad_h2[e];
        if ((word & 0x7fff7fffu) == 0u) continue;
        a.grad_h2[e] = 0u;
        const float g[2] = {half_lo(word) * unscale, half_hi(word) * unscale};
        if (!isfinite_(g[0]) || !isfinite_(g[1])) continue;
        f4* rec = reinterpret_cast<f4*>(a.state + 8 * (size_t)e);
        f4 pm = rec[0], vv = rec[1];
        uint16_t bits[2];
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const float m = a.beta1 * pm[2 + k] + (1.0f - a.beta1) * g[k];
            const float v = a.beta2 * vv[k] + (1.0f - a.beta2) * (g[k] * g[k]);
            const float mh = m / a.bias1;
            const float vh = v / a.bias2;
            const float p = pm[k] - a.lr * mh / (sqrt_(vh) + a.eps);
            pm[k] = p;
            pm[2 + k] = m;
            vv[k] = v;
            const _Float16 h = (_Float16)p;
            __builtin_memcpy(&bits[k], &h, 2);
        }
        rec[0] = pm;
        rec[1] = vv;
        a.half_out[e] = (uint32_t)bits[0] | ((uint32_t)bits[1] << 16);
    }
}

// Master table values in or out of the Adam records (the moments stay);
// setting them also writes the FP16 copy.
__global__ void rc_table_master_kernel(float* state, float* master, uint32_t* table_h, size_t pairs, int set) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < pairs; e += stride) {
        float* rec = state + 8 * e;
        if (set) {
            rec[0] = master[2 * e];
            rec[1] = master[2 * e + 1];
            const _Float16 h0 = (_Float16)rec[0], h1 = (_Float16)rec[1];
            uint16_t b0, b1;
            __builtin_memcpy(&b0, &h0, 2);
            __builtin_memcpy(&b1, &h1, 2);
            table_h[e] = (uint32_t)b0 | ((uint32_t)b1 << 16);
        } else {
            master[2 * e] = rec[0];
            master[2 * e + 1] = rec[1];
        }
    }
}

__global__ void rc_targets_kernel(RcTrainQueue q, const float* slot_radiance, uint32_t c, uint32_t capacity,
                                  float* target) {
    const uint32_t count = c < capacity ? c : capacity;
    const float nan = __builtin_nanf("");
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < count; r += gridDim.x * blockDim.x) {
        const uint32_t s = q.slot[r];
        const float t[3] = {q.tx[r], q.ty[r], q.tz[r]};
        const float rr[3] = {q.rx[r], q.ry[r], q.rz[r]};
        for (int k = 0; k < 3; k++) {
            const float v = (slot_radiance[3 * (size_t)s + k] - rr[k]) / t[k];
            target[(size_t)k * capacity + r] = (t[k] > 1e-8f && isfinite_(v)) ? v : nan;
        }
    }
}

__global__ void rc_fill_kernel(float* p, float v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void rc_halves_kernel(const float* mlp, uint16_t* mlp_bits) {
    _Float16* mlp_h = reinterpret_cast<_Float16*>(mlp_bits);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)kRcMlpParams; i += gridDim.x * blockDim.x)
        mlp_h[i] = (_Float16)mlp[i];
}

uint32_t rc_train_waves(uint32_t batch) {
    return ((batch + 32u * kTrainWaves - 1u) / (32u * kTrainWaves)) * kTrainWaves;
}

hipError_t launch_rc_infer(const RcInferArgs& a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(rc_infer_kernel, dim3(blocks), dim3(64 * kInferWaves), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_rc_sort_queries(const RcSortArgs& a, int blocks, hipStream_t s) {
    hipLaunchKernelGGL(rc_qkeys_kernel, dim3(blocks), dim3(kRcSortBlock), 0, s, a);
    hipLaunchKernelGGL(rc_qscan_kernel, dim3(1), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(rc_qscatter_kernel, dim3(blocks), dim3(kRcSortBlock), 0, s, a);
    hipLaunchKernelGGL(rc_qclear_kernel, dim3(16), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_rc_train(const RcParams& p, const RcSamples& smp, uint32_t batch, const RcGrads& g, float scale,
                           hipStream_t s) {
    const uint32_t blocks = rc_train_waves(batch) / kTrainWaves;
    hipLaunchKernelGGL(rc_train_kernel, dim3(blocks), dim3(64 * kTrainWaves), 0, s, p, smp, batch, g, scale);
    return hipGetLastError();
}
hipError_t launch_rc_reduce(const float* partial, uint32_t waves, float* grad_mlp, hipStream_t s) {
    const dim3 grid((kRcMlpParams + 255) / 256, (waves + kReduceSlice - 1) / kReduceSlice);
    hipLaunchKernelGGL(rc_reduce_kernel, grid, dim3(256), 0, s, partial, waves, grad_mlp);
    return hipGetLastError();
}
hipError_t launch_rc_adam_mlp(const RcAdam& a, hipStream_t s) {
    hipLaunchKernelGGL(rc_adam_mlp_kernel, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_rc_adam_table(const RcAdam& a, hipStream_t s) {
    hipLaunchKernelGGL(rc_adam_table_kernel, dim3(2048), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_rc_targets(const RcTrainQueue& q, const float* slot_radiance, uint32_t count, uint32_t capacity,
                             float* target, hipStream_t s) {
    hipLaunchKernelGGL(rc_targets_kernel, dim3(1024), dim3(256), 0, s, q, slot_radiance, count, capacity, target);
    return hipGetLastError();
}
hipError_t launch_rc_fill(float* p, float v, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(rc_fill_kernel, dim3(1024), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

}  // namespace rm

namespace rm {
hipError_t launch_rc_halves(const float* mlp, uint16_t* mlp_h, hipStream_t s) {
    hipLaunchKernelGGL(rc_halves_kernel, dim3(36), dim3(256), 0, s, mlp, mlp_h);
    return hipGetLastError();
}
hipError_t launch_rc_table_master(float* state, float* master, uint32_t* table_h, size_t table_floats, bool set,
                                  hipStream_t s) {
    hipLaunchKernelGGL(rc_table_master_kernel, dim3(2048), dim3(256), 0, s, state, master, table_h, table_floats / 2,
                       set ? 1 : 0);
    return hipGetLastError();
}
}  // namespace rm
