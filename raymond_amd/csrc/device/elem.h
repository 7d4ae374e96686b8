// Elementary functions with identical results on host and device.
//
// Each float function evaluates in double precision from +, -, *, / and
// sqrt only (all correctly rounded on both x86-64 SSE2 and gfx950), then
// rounds once to float.  Methods (J.-M. Muller, "Elementary Functions:
// Algorithms and Implementation", chapters on range reduction and
// polynomial approximation):
//   sin/cos  Cody-Waite reduction by pi/2 (three-part constant), Taylor
//            series to degree 17 on |r| <= pi/4.
//   atan     reduction to |t| <= tan(pi/8) by atan(t) = pi/4 + atan((t-1)/(t+1)),
//            one argument halving atan(z) = 2 atan(z / (1 + sqrt(1 + z^2))),
//            odd Taylor series to degree 25.
//   log      x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
//            s = (m-1)/(m+1), odd series to degree 23.
//   exp      x = k ln2 + r (two-part ln2), Taylor series to degree 15, scale
//            by 2^k through the exponent field.
// The double result is within a few 1e-16 relative of the true value, so the
// final float rounding is correct except in rare near-halfway cases; the
// host/device agreement is exact by construction.
#pragma once
#include "common.h"

namespace rm {

constexpr double kPio2Hi = 0x1.921fb54000000p+0;
constexpr double kPio2Mid = 0x1.10b4611800000p-30;
constexpr double kPio2Lo = 0x1.313198a2e0370p-61;
constexpr double kTwoOverPi = 0.6366197723675814;
constexpr double kPiD = 3.141592653589793;
constexpr double kPio2D = 1.5707963267948966;
constexpr double kPio4D = 0.7853981633974483;
constexpr double kLn2Hi = 0x1.62e42fee00000p-1;
constexpr double kLn2Lo = 1.9082149292705877e-10;
constexpr double kLn2D = 0.6931471805599453;
constexpr double kInvLn2D = 1.4426950408889634;
constexpr double kSqrt2D = 1.4142135623730951;
constexpr double kTanPi8D = 0.41421356237309503;

// sin(r), cos(r) for |r| <= pi/4 (Taylor, Horner form)
RM_HD double sin_poly(double r) {
    const double r2 = r * r;
    double p = 1.0 / 355687428096000.0;          // 1/17!
    p = p * r2 - 1.0 / 1307674368000.0;          // 1/15!
    p = p * r2 + 1.0 / 6227020800.0;             // 1/13!
    p = p * r2 - 1.0 / 39916800.0;               // 1/11!
    p = p * r2 + 1.0 / 362880.0;                 // 1/9!
    p = p * r2 - 1.0 / 5040.0;                   // 1/7!
    p = p * r2 + 1.0 / 120.0;                    // 1/5!
    p = p * r2 - 1.0 / 6.0;                      // 1/3!
    return r + r * (r2 * p);
}
RM_HD double cos_poly(double r) {
    const double r2 = r * r;
    double p = 1.0 / 6402373705728000.0;         // 1/18!
    p = p * r2 - 1.0 / 20922789888000.0;         // 1/16!
    p = p * r2 + 1.0 / 87178291200.0;            // 1/14!
    p = p * r2 - 1.0 / 479001600.0;              // 1/12!
    p = p * r2 + 1.0 / 3628800.0;                // 1/10!
    p = p * r2 - 1.0 / 40320.0;                  // 1/8!
    p = p * r2 + 1.0 / 720.0;                    // 1/6!
    p = p * r2 - 1.0 / 24.0;                     // 1/4!
    p = p * r2 + 0.5;                            // 1/2!
    return 1.0 - r2 * p;
}

// Reduce x by multiples of pi/2: returns r in [-pi/4, pi/4] and quadrant k mod 4.
RM_HD double reduce_pio2(double x, int& quadrant) {
    const double k = drint_(x * kTwoOverPi);
    double r = x - k * kPio2Hi;
    r = r - k * kPio2Mid;
    r = r - k * kPio2Lo;
    // k is an integer < 2^53; take it modulo 4 without overflow concerns for
    // the float range used here (|x| < 2^31 handled, beyond that precision
    // of the argument itself is meaningless).
    const double k4 = k - 4.0 * dfloor_(k * 0.25);
    quadrant = (int)k4;
    return r;
}

RM_HD void sincos_d(double x, double& s, double& c) {
    int q;
    const double r = reduce_pio2(x, q);
    const double sr = sin_poly(r), cr = cos_poly(r);
    switch (q & 3) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
    }
}

RM_HD float sin_(float x) {
    if (!isfinite_(x)) return u2f(0x7fc00000u);
    double s, c;
    sincos_d((double)x, s, c);
    return (float)s;
}
RM_HD float cos_(float x) {
    if (!isfinite_(x)) return u2f(0x7fc00000u);
    double s, c;
    sincos_d((double)x, s, c);
    return (float)c;
}
// Metal-style sincos: returns sin, writes cos.
RM_HD float sincos_(float x, float& c) {
    if (!isfinite_(x)) { c = u2f(0x7fc00000u); return c; }
    double sd, cd;
    sincos_d((double)x, sd, cd);
    c = (float)cd;
    return (float)sd;
}
RM_HD float tan_(float x) {
    if (!isfinite_(x)) return u2f(0x7fc00000u);
    double s, c;
    sincos_d((double)x, s, c);
    return (float)(s / c);
}

// atan for t in [0, 1]
RM_HD double atan01_d(double t) {
    double base = 0.0;
    if (t > kTanPi8D) {
        base = kPio4D;
        t = (t - 1.0) / (t + 1.0);
    }
    // halve the argument: atan(t) = 2 atan(t / (1 + sqrt(1 + t^2)))
    const double h = t / (1.0 + dsqrt_(1.0 + t * t));
    const double h2 = h * h;
    double p = 1.0 / 25.0;
    p = 1.0 / 23.0 - h2 * p;
    p = 1.0 / 21.0 - h2 * p;
    p = 1.0 / 19.0 - h2 * p;
    p = 1.0 / 17.0 - h2 * p;
    p = 1.0 / 15.0 - h2 * p;
    p = 1.0 / 13.0 - h2 * p;
    p = 1.0 / 11.0 - h2 * p;
    p = 1.0 / 9.0 - h2 * p;
    p = 1.0 / 7.0 - h2 * p;
    p = 1.0 / 5.0 - h2 * p;
    p = 1.0 / 3.0 - h2 * p;
    const double a = h - h * (h2 * p);
    return base + 2.0 * a;
}

RM_HD double atan2_d(double y, double x) {
    const double ax = dabs_(x), ay = dabs_(y);
    const bool xneg = (d2u(x) >> 63) != 0;
    double a;
    if (ay == 0.0) {
        a = xneg ? kPiD : 0.0;
    } else if (ax == 0.0) {
        a = kPio2D;
    } else if (ax == __builtin_huge_val() && ay == __builtin_huge_val()) {
        a = xneg ? 3.0 * kPio4D : kPio4D;
    } else if (ax == __builtin_huge_val()) {
        a = xneg ? kPiD : 0.0;
    } else if (ay == __builtin_huge_val()) {
        a = kPio2D;
    } else {
        if (ay > ax) {
            a = kPio2D - atan01_d(ax / ay);
        } else {
            a = atan01_d(ay / ax);
        }
        if (xneg) a = kPiD - a;
    }
    return (d2u(y) >> 63) ? -a : a;
}

RM_HD float atan2_(float y, float x) {
    if (isnan_(x) || isnan_(y)) return u2f(0x7fc00000u);
    return (float)atan2_d((double)y, (double)x);
}
RM_HD float atan_(float x) {
    if (isnan_(x)) return x;
    return (float)atan2_d((double)x, 1.0);
}
RM_HD float acos_(float x) {
    if (!(x >= -1.0f && x <= 1.0f)) return u2f(0x7fc00000u);
    const double xd = (double)x;
    const double s = dsqrt_((1.0 - xd) * (1.0 + xd));
    return (float)atan2_d(s, xd);
}
RM_HD float asin_(float x) {
    if (!(x >= -1.0f && x <= 1.0f)) return u2f(0x7fc00000u);
    const double xd = (double)x;
    const double c = dsqrt_((1.0 - xd) * (1.0 + xd));
    return (float)atan2_d(xd, c);
}

// natural log for finite positive normal double input
RM_HD double log_pos_d(double x) {
    uint64_t bits = d2u(x);
    int e = (int)((bits >> 52) & 0x7ff);
    if (e == 0) {  // subnormal double: scale up
        x = x * 18014398509481984.0;  // 2^54
        bits = d2u(x);
        e = (int)((bits >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = u2d((bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
    if (m > kSqrt2D) { m = m * 0.5; e += 1; }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    const double lm = 2.0 * (s + s * (s2 * p));
    const double ed = (double)e;
    return (ed * kLn2Hi + lm) + ed * kLn2Lo;
}

RM_HD float log_(float x) {
    if (isnan_(x)) return x;
    if (x < 0.0f) return u2f(0x7fc00000u);
    if (x == 0.0f) return -kInf;
    if (isinf_(x)) return x;
    return (float)log_pos_d((double)x);
}
RM_HD float log2_(float x) {
    if (isnan_(x)) return x;
    if (x < 0.0f) return u2f(0x7fc00000u);
    if (x == 0.0f) return -kInf;
    if (isinf_(x)) return x;
    return (float)(log_pos_d((double)x) * kInvLn2D);
}

// exp in double for |x| < ~700 (callers clamp)
RM_HD double exp_d(double x) {
    const double k = drint_(x * kInvLn2D);
    double r = x - k * kLn2Hi;
    r = r - k * kLn2Lo;
    double p = 1.0 / 1307674368000.0;  // 1/15!
    p = p * r + 1.0 / 87178291200.0;   // 1/14!
    p = p * r + 1.0 / 6227020800.0;    // 1/13!
    p = p * r + 1.0 / 479001600.0;     // 1/12!
    p = p * r + 1.0 / 39916800.0;      // 1/11!
    p = p * r + 1.0 / 3628800.0;       // 1/10!
    p = p * r + 1.0 / 362880.0;        // 1/9!
    p = p * r + 1.0 / 40320.0;         // 1/8!
    p = p * r + 1.0 / 5040.0;          // 1/7!
    p = p * r + 1.0 / 720.0;           // 1/6!
    p = p * r + 1.0 / 120.0;           // 1/5!
    p = p * r + 1.0 / 24.0;            // 1/4!
    p = p * r + 1.0 / 6.0;             // 1/3!
    p = p * r + 0.5;                   // 1/2!
    p = p * r + 1.0;
    p = p * r + 1.0;
    // scale by 2^k in two steps so that large |k| neither overflows the
    // exponent field nor loses subnormal results.
    const int ki = (int)k;
    const int k1 = ki / 2, k2 = ki - k1;
    const double s1 = u2d((uint64_t)(k1 + 1023) << 52);
    const double s2 = u2d((uint64_t)(k2 + 1023) << 52);
    return (p * s1) * s2;
}

RM_HD float exp_(float x) {
    if (isnan_(x)) return x;
    if (x > 89.0f) return kInf;
    if (x < -104.0f) return 0.0f;
    return (float)exp_d((double)x);
}
RM_HD float exp2_(float x) {
    if (isnan_(x)) return x;
    if (x > 128.5f) return kInf;
    if (x < -150.5f) return 0.0f;
    return (float)exp_d((double)x * kLn2D);
}

RM_HD bool is_int_(float y) { return floor_(y) == y; }
RM_HD bool is_odd_int_(float y) {
    if (!is_int_(y) || abs_(y) >= 16777216.0f) return false;
    const int64_t i = (int64_t)y;
    return (i & 1) != 0;
}

// pow with the C99 special cases (Annex F.9.4.4)
RM_HD float pow_(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x == 1.0f) return 1.0f;
    if (isnan_(x) || isnan_(y)) return u2f(0x7fc00000u);
    const float ax = abs_(x);
    if (isinf_(y)) {
        if (ax == 1.0f) return 1.0f;
        const bool big = ax > 1.0f;
        return (big == (y > 0.0f)) ? kInf : 0.0f;
    }
    const bool yodd = is_odd_int_(y);
    if (x == 0.0f) {
        if (y < 0.0f) return yodd ? copysign_(kInf, x) : kInf;
        return yodd ? x : 0.0f;
    }
    if (isinf_(x)) {
        if (x < 0.0f) {
            if (y < 0.0f) return yodd ? -0.0f : 0.0f;
            return yodd ? -kInf : kInf;
        }
        return y < 0.0f ? 0.0f : kInf;
    }
    float sgn = 1.0f;
    if (x < 0.0f) {
        if (!is_int_(y)) return u2f(0x7fc00000u);
        if (yodd) sgn = -1.0f;
    }
    const double l = (double)y * log_pos_d((double)ax);
    float r;
    if (l > 89.0) r = kInf;
    else if (l < -104.0) r = 0.0f;
    else r = (float)exp_d(l);
    return sgn * r;
}

}  // namespace rm
