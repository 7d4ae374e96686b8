// Radiance-cache oracle -- TEST INFRASTRUCTURE ONLY.
//
// Double-precision restatement of the cache's forward pass, relative-L2 loss
// and backward pass (hash-grid encoding, 48-64-64-3 ReLU MLP without biases)
// for checking the GPU kernels.  Parameters enter as the FP32 masters and are
// rounded to FP16 here, as the GPU kernels read them; everything after that
// is double.  The grid geometry (cell lookup, dense/hash index) comes from
// raymond_amd/csrc/device/rcache.h, which tests/test_rcache_cpu.py pins
// against an independent numpy restatement.  No reference counterpart: the
// cache is not part of the reference renderer.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../raymond_amd/csrc/device/rcache.h"

using namespace rm;

namespace {

// float -> binary16 -> float with round-to-nearest-even, subnormals included.
float round_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = x & 0x80000000u;
    const uint32_t ax = x & 0x7fffffffu;
    uint32_t h;
    if (ax >= 0x7f800000u) {
        h = ax > 0x7f800000u ? 0x7e00u : 0x7c00u;  // NaN, inf
    } else if (ax >= 0x477ff000u) {
        h = 0x7c00u;  // rounds to inf (>= 65520)
    } else if (ax < 0x38800000u) {
        // subnormal half: value = m * 2^-24
        const float a = std::fabs(f) * 16777216.0f;  // exact scaling by 2^24
        h = (uint32_t)std::nearbyint(a);             // default rounding mode: ties to even
    } else {
        const uint32_t e = (ax >> 23) - 112u;  // rebias 127 -> 15
        const uint32_t m = ax & 0x7fffffu;
        uint32_t hm = m >> 13;
        const uint32_t rest = m & 0x1fffu;
        uint32_t he = e;
        if (rest > 0x1000u || (rest == 0x1000u && (hm & 1u))) {
            hm++;
            if (hm == 0x400u) {
                hm = 0;
                he++;
            }
        }
        h = (he << 10) | hm;
    }
    // decode
    const uint32_t he = (h >> 10) & 0x1fu, hm = h & 0x3ffu;
    float out;
    if (he == 0) out = std::ldexp((float)hm, -24);
    else if (he == 31) out = hm ? NAN : INFINITY;
    else out = std::ldexp((float)(hm | 0x400u), (int)he - 25);
    return sign ? -out : out;
}

struct Grid {
    RcGrid g;
    size_t T;
};

Grid make_grid(uint32_t log2_table) {
    Grid G{};
    G.g.log2_table = log2_table;
    G.T = (size_t)1 << log2_table;
    const double b = std::exp((std::log(2048.0) - std::log(16.0)) / (kRcLevels - 1));
    for (int l = 0; l < kRcLevels; l++) {
        const uint32_t res = (uint32_t)std::floor(16.0 * std::pow(b, l) + 0.5);
        G.g.res[l] = res;
        const uint64_t n = res + 1ull;
        G.g.dense[l] = n * n * n <= G.T ? 1u : 0u;
    }
    return G;
}

struct Corner {
    uint32_t idx[8];
    double w[8];
};

Corner corners(const Grid& G, int level, const float* pos) {
    Corner c;
    uint32_t cell[3];
    float frac[3];
    for (int a = 0; a < 3; a++) rc_cell(pos[a], G.g.res[level], cell[a], frac[a]);
    for (int k = 0; k < 8; k++) {
        const uint32_t b[3] = {(uint32_t)(k & 1), (uint32_t)((k >> 1) & 1), (uint32_t)((k >> 2) & 1)};
        c.idx[k] = rc_index(G.g, level, cell[0] + b[0], cell[1] + b[1], cell[2] + b[2]);
        double w = 1.0;
        for (int a = 0; a < 3; a++) w *= b[a] ? (double)frac[a] : 1.0 - (double)frac[a];
        c.w[k] = w;
    }
    return c;
}

struct Net {
    std::vector<double> w1, w2, w3;  // [in][out]
    std::vector<double> table;       // FP16-rounded
};

Net load(const Grid& G, const float* mlp, const float* table) {
    Net n;
    n.w1.resize(kRcW1);
    n.w2.resize(kRcW2);
    n.w3.resize(kRcW3);
    for (int i = 0; i < kRcW1; i++) n.w1[i] = round_half(mlp[i]);
    for (int i = 0; i < kRcW2; i++) n.w2[i] = round_half(mlp[kRcW1 + i]);
    for (int i = 0; i < kRcW3; i++) n.w3[i] = round_half(mlp[kRcW1 + kRcW2 + i]);
    const size_t tf = (size_t)kRcLevels * G.T * 2;
    n.table.resize(tf);
    for (size_t i = 0; i < tf; i++) n.table[i] = round_half(table[i]);
    return n;
}

struct Fwd {
    double x[kRcInputs], z1[kRcHidden], h1[kRcHidden], z2[kRcHidden], h2[kRcHidden], y[kRcOutputs];
    Corner c[kRcLevels];
};

void forward(const Grid& G, const Net& n, const float* f, Fwd& r) {
    for (int l = 0; l < kRcLevels; l++) {
        r.c[l] = corners(G, l, f);
        const double* tab = n.table.data() + (size_t)l * G.T * 2;
        double f0 = 0.0, f1 = 0.0;
        for (int k = 0; k < 8; k++) {
            f0 += r.c[l].w[k] * tab[2 * (size_t)r.c[l].idx[k]];
            f1 += r.c[l].w[k] * tab[2 * (size_t)r.c[l].idx[k] + 1];
        }
        r.x[2 * l] = f0;
        r.x[2 * l + 1] = f1;
    }
    for (int e = 0; e < kRcExtra; e++) r.x[kRcEncoded + e] = f[3 + e];
    r.x[kRcEncoded + kRcExtra] = 1.0;
    for (int i = kRcEncoded + kRcExtra + 1; i < kRcInputs; i++) r.x[i] = 0.0;
    for (int o = 0; o < kRcHidden; o++) {
        double s = 0.0;
        for (int i = 0; i < kRcInputs; i++) s += r.x[i] * n.w1[(size_t)i * kRcHidden + o];
        r.z1[o] = s;
        r.h1[o] = s > 0.0 ? s : 0.0;
    }
    for (int o = 0; o < kRcHidden; o++) {
        double s = 0.0;
        for (int i = 0; i < kRcHidden; i++) s += r.h1[i] * n.w2[(size_t)i * kRcHidden + o];
        r.z2[o] = s;
        r.h2[o] = s > 0.0 ? s : 0.0;
    }
    for (int o = 0; o < kRcOutputs; o++) {
        double s = 0.0;
        for (int i = 0; i < kRcHidden; i++) s += r.h2[i] * n.w3[(size_t)i * kRcW3Cols + o];
        r.y[o] = s;
    }
}

}  // namespace

extern "C" {

float or_rc_round_half(float f) { return round_half(f); }

// The float atan2 / acos of the vertex features (rcache.h), for the tests.
float or_rc_atan2(float y, float x) { return rm::rc_atan2(y, x); }
float or_rc_acos(float z) { return rm::rc_acos(z); }

void or_rc_grid(uint32_t log2_table, uint32_t* res, uint32_t* dense) {
    const Grid G = make_grid(log2_table);
    for (int l = 0; l < kRcLevels; l++) {
        res[l] = G.g.res[l];
        dense[l] = G.g.dense[l];
    }
}

// Table index of grid point (x, y, z) at `level` (rcache.h's rule).
uint32_t or_rc_index(uint32_t log2_table, int level, uint32_t x, uint32_t y, uint32_t z) {
    const Grid G = make_grid(log2_table);
    return rc_index(G.g, level, x, y, z);
}

// n x 12 features -> n x 3 outputs.
void or_rc_forward(uint32_t log2_table, const float* mlp, const float* table, const float* features, uint32_t n,
                   double* out) {
    const Grid G = make_grid(log2_table);
    const Net net = load(G, mlp, table);
    Fwd r;
    for (uint32_t s = 0; s < n; s++) {
        forward(G, net, features + (size_t)s * kRcFeatureFloats, r);
        for (int c = 0; c < 3; c++) out[(size_t)s * 3 + c] = r.y[c];
    }
}

// Gradients of the mean relative-L2 loss over the batch.  grad_mlp has
// kRcMlpParams entries ([in][out] per layer, W3 with 32 columns), grad_table
// the table's levels x 2^log2 x 2.  Samples whose target is not finite are
// skipped (they still count in the mean, as on the GPU).
void or_rc_gradients(uint32_t log2_table, const float* mlp, const float* table, const float* features,
                     const float* targets, uint32_t n, double* grad_mlp, double* grad_table, double* loss) {
    const Grid G = make_grid(log2_table);
    const Net net = load(G, mlp, table);
    std::memset(grad_mlp, 0, sizeof(double) * kRcMlpParams);
    std::memset(grad_table, 0, sizeof(double) * (size_t)kRcLevels * G.T * 2);
    double L = 0.0;
    const double inv_n = 1.0 / (double)n;
    Fwd r;
    for (uint32_t s = 0; s < n; s++) {
        const float* t = targets + (size_t)s * 3;
        if (!std::isfinite(t[0]) || !std::isfinite(t[1]) || !std::isfinite(t[2])) continue;
        forward(G, net, features + (size_t)s * kRcFeatureFloats, r);
        double dy[kRcOutputs];
        for (int c = 0; c < 3; c++) {
            const double den = r.y[c] * r.y[c] + 0.01;
            const double d = r.y[c] - (double)t[c];
            L += d * d / den * inv_n;
            dy[c] = 2.0 * d / den * inv_n;
        }
        double dz2[kRcHidden], dz1[kRcHidden];
        for (int i = 0; i < kRcHidden; i++) {
            double s3 = 0.0;
            for (int c = 0; c < 3; c++) {
                grad_mlp[kRcW1 + kRcW2 + (size_t)i * kRcW3Cols + c] += r.h2[i] * dy[c];
                s3 += net.w3[(size_t)i * kRcW3Cols + c] * dy[c];
            }
            dz2[i] = r.z2[i] > 0.0 ? s3 : 0.0;
        }
        for (int i = 0; i < kRcHidden; i++) {
            double s2 = 0.0;
            for (int o = 0; o < kRcHidden; o++) {
                grad_mlp[kRcW1 + (size_t)i * kRcHidden + o] += r.h1[i] * dz2[o];
                s2 += net.w2[(size_t)i * kRcHidden + o] * dz2[o];
            }
            dz1[i] = r.z1[i] > 0.0 ? s2 : 0.0;
        }
        for (int i = 0; i < kRcInputs; i++)
            for (int o = 0; o < kRcHidden; o++) grad_mlp[(size_t)i * kRcHidden + o] += r.x[i] * dz1[o];
        for (int l = 0; l < kRcLevels; l++) {
            double g[2] = {0.0, 0.0};
            for (int f = 0; f < 2; f++)
                for (int o = 0; o < kRcHidden; o++) g[f] += net.w1[(size_t)(2 * l + f) * kRcHidden + o] * dz1[o];
            double* gt = grad_table + (size_t)l * G.T * 2;
            for (int k = 0; k < 8; k++) {
                gt[2 * (size_t)r.c[l].idx[k]] += r.c[l].w[k] * g[0];
                gt[2 * (size_t)r.c[l].idx[k] + 1] += r.c[l].w[k] * g[1];
            }
        }
    }
    *loss = L;
}

}  // extern "C"
