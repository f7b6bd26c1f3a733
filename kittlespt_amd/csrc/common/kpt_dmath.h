// kpt_dmath.h — deterministic transcendental functions for host (g++) and
// device (hipcc, gfx950).
//
// Why this exists: the path tracer's control flow depends on float values
// (Russian roulette, lobe choice, sun-vs-area NEE choice), so a one-ulp
// difference between the host C library and the device math library would send
// a path down a different branch.  Every function here is built only from
// IEEE-754 operations that are correctly rounded on both targets (+ - * /,
// sqrt, rint, floor, bit casts, int<->float conversions), evaluated in a fixed
// order; compiled with -ffp-contract=off (no FMA contraction) on both sides
// they return bit-identical results on the CPU and on the GPU.
//
// Methods (all textbook, coefficients are exact Taylor/series coefficients
// folded at compile time, not fitted constants of any vendor library):
//   expf    : Cody–Waite reduction x = k·ln2 + r (ln2 split hi/lo so k·hi is
//             exact), degree-7 Taylor series of e^r on |r| <= ln2/2, scale 2^k.
//   exp2f   : x = k + r, e^(r·ln2) by its Taylor series, scale 2^k.
//   log2f   : x = m·2^e, m in [1/sqrt2, sqrt2), ln m = 2·atanh(s),
//             s = (m-1)/(m+1), odd series of atanh, evaluated in double.
//   powf    : exp2(y·log2 x) with log2/exp2 carried in double precision, one
//             final rounding to float; C99 special cases handled explicitly.
//   sinf/cosf/tanf : reduction by pi/2 in double (pi/2 split into a 33-bit
//             head and a tail), Taylor series of sin/cos on |r| <= pi/4.
//   acosf   : acos x = pi/2 - asin x; asin by its power series on |x| <= 1/2
//             and asin x = pi/2 - 2·asin(sqrt((1-x)/2)) above, in double.
// Accuracy against glibc is checked in tests/test_oracle.py (test_dmath_*).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define KPT_HD __host__ __device__ inline
#else
#define KPT_HD inline
#endif

namespace kpt {
namespace dm {

KPT_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
KPT_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
KPT_HD uint64_t d2u(double d) { return __builtin_bit_cast(uint64_t, d); }
KPT_HD double u2d(uint64_t u) { return __builtin_bit_cast(double, u); }

KPT_HD float f_inf() { return u2f(0x7f800000u); }
KPT_HD float f_nan() { return u2f(0x7fc00000u); }
KPT_HD bool f_isnan(float x) { return x != x; }
KPT_HD bool f_isinf(float x) { return (f2u(x) & 0x7fffffffu) == 0x7f800000u; }

// 2^n as a float / double, n inside the normal exponent range.
KPT_HD float pow2f(int n) { return u2f(uint32_t(n + 127) << 23); }
KPT_HD double pow2d(int n) { return u2d(uint64_t(n + 1023) << 52); }

// x * 2^n for n in [-300, 300] using exact power-of-two multiplies.
KPT_HD float scalef(float x, int n) {
    if (n > 127) {
        x *= pow2f(127); n -= 127;
        if (n > 127) { x *= pow2f(127); n -= 127; if (n > 127) n = 127; }
    } else if (n < -126) {
        x *= pow2f(-126); n += 126;
        if (n < -126) { x *= pow2f(-126); n += 126; if (n < -126) n = -126; }
    }
    return x * pow2f(n);
}

KPT_HD double scaled(double x, int n) {
    if (n > 1023) { x *= pow2d(1023); n -= 1023; if (n > 1023) n = 1023; }
    else if (n < -1022) { x *= pow2d(-1022); n += 1022; if (n < -1022) n = -1022; }
    return x * pow2d(n);
}

// ---------------------------------------------------------------- exp family
// ln 2 = LN2_HI + LN2_LO, LN2_HI has 15 significant bits so k*LN2_HI is exact
// for |k| < 512.
#define KPT_LN2_HI 0.693145751953125f
#define KPT_LN2_LO 1.42860682030941723212e-6f
#define KPT_LOG2E 1.44269504088896340736f

KPT_HD float expf(float x) {
    if (f_isnan(x)) return x;
    if (x > 88.7228394f) return f_inf();
    if (x < -103.972084f) return 0.0f;
    float k = __builtin_rintf(x * KPT_LOG2E);
    float r = (x - k * KPT_LN2_HI) - k * KPT_LN2_LO;
    float p = (float)(1.0 / 5040.0);
    p = p * r + (float)(1.0 / 720.0);
    p = p * r + (float)(1.0 / 120.0);
    p = p * r + (float)(1.0 / 24.0);
    p = p * r + (float)(1.0 / 6.0);
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return scalef(p, (int)k);
}

KPT_HD float exp2f(float x) {
    if (f_isnan(x)) return x;
    if (x >= 128.0f) return f_inf();
    if (x < -150.0f) return 0.0f;
    float k = __builtin_rintf(x);
    float r = (x - k) * 0.693147180559945309f;  // |r| <= ln2/2
    float p = (float)(1.0 / 5040.0);
    p = p * r + (float)(1.0 / 720.0);
    p = p * r + (float)(1.0 / 120.0);
    p = p * r + (float)(1.0 / 24.0);
    p = p * r + (float)(1.0 / 6.0);
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return scalef(p, (int)k);
}

// 2^t in double for |t| < 1100, degree-13 Taylor of e^(r ln2), |r| <= 1/2.
KPT_HD double exp2_d(double t) {
    double k = __builtin_rint(t);
    double r = (t - k) * 0.69314718055994530942;
    double p = 1.0 / 6227020800.0;            // 1/13!
    p = p * r + 1.0 / 479001600.0;            // 1/12!
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return scaled(p, (int)k);
}

// ---------------------------------------------------------------- log family
// Splits a positive finite float into m in [1/sqrt2, sqrt2) and exponent e.
KPT_HD void split_mant(float x, float* m, int* e) {
    uint32_t u = f2u(x);
    int ee = 0;
    if (u < 0x00800000u) {  // subnormal: scale into the normal range exactly
        x *= 8388608.0f;
        u = f2u(x);
        ee = -23;
    }
    ee += int(u >> 23) - 127;
    float mm = u2f((u & 0x007fffffu) | 0x3f800000u);
    if (mm > 1.41421356f) { mm *= 0.5f; ee += 1; }
    *m = mm;
    *e = ee;
}

// log2 of a positive finite float carried in double (relative error ~1e-16).
KPT_HD double log2_pos_d(float x) {
    float m; int e;
    split_mant(x, &m, &e);
    double md = (double)m;
    double s = (md - 1.0) / (md + 1.0);
    double s2 = s * s;
    // atanh(s)/s = sum s^(2k)/(2k+1), |s| <= 0.1716 -> 11 terms reach 1e-18
    double p = 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    double ln_m = 2.0 * s * p;
    return (double)e + ln_m * 1.44269504088896340736;
}

KPT_HD float log2f(float x) {
    if (f_isnan(x)) return x;
    if (x < 0.0f) return f_nan();
    if (x == 0.0f) return -f_inf();
    if (f_isinf(x)) return x;
    return (float)log2_pos_d(x);
}

KPT_HD float logf(float x) {
    if (f_isnan(x)) return x;
    if (x < 0.0f) return f_nan();
    if (x == 0.0f) return -f_inf();
    if (f_isinf(x)) return x;
    return (float)(log2_pos_d(x) * 0.69314718055994530942);
}

// true when the float y is an integer; *odd set when it is an odd integer
KPT_HD bool f_isint(float y, bool* odd) {
    float t = __builtin_truncf(y);
    if (t != y) { *odd = false; return false; }
    // |y| >= 2^24 are all even integers
    if (__builtin_fabsf(y) >= 16777216.0f) { *odd = false; return true; }
    int32_t i = (int32_t)y;
    *odd = (i & 1) != 0;
    return true;
}

KPT_HD float powf(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x == 1.0f) return 1.0f;
    if (f_isnan(x) || f_isnan(y)) return f_nan();
    if (y == 1.0f) return x;
    bool odd = false;
    bool yint = f_isint(y, &odd);
    float sign = 1.0f;
    if (x < 0.0f) {
        if (f_isinf(x)) {
            if (y < 0.0f) return (yint && odd) ? -0.0f : 0.0f;
            return (yint && odd) ? -f_inf() : f_inf();
        }
        if (!yint) return f_nan();
        if (odd) sign = -1.0f;
        x = -x;
    }
    if (x == 0.0f) {
        bool neg0 = (f2u(x) >> 31) != 0 || sign < 0.0f;
        if (y > 0.0f) return (odd && neg0) ? -0.0f : 0.0f;
        return (odd && neg0) ? -f_inf() : f_inf();
    }
    if (f_isinf(x)) return (y > 0.0f) ? f_inf() : 0.0f;
    if (f_isinf(y)) {
        if (x == 1.0f) return 1.0f;
        bool big = x > 1.0f;
        if (y > 0.0f) return big ? f_inf() : 0.0f;
        return big ? 0.0f : f_inf();
    }
    double t = (double)y * log2_pos_d(x);
    if (t >= 128.0) return sign * f_inf();
    if (t < -152.0) return sign * 0.0f;
    return sign * (float)exp2_d(t);
}

// ---------------------------------------------------------------- trig family
// pi/2 = PIO2_HI + PIO2_LO with PIO2_HI holding 33 significant bits, so k*HI is
// exact in double for |k| < 2^20.
#define KPT_PIO2_HI 1.57079632673412561417e+00
#define KPT_PIO2_LO 6.07710050650619224932e-11
#define KPT_TWO_OVER_PI 6.36619772367581382433e-01

KPT_HD double sin_poly(double r) {
    double r2 = r * r;
    double p = -1.0 / 1307674368000.0;     // -1/15!
    p = p * r2 + 1.0 / 6227020800.0;       // 1/13!
    p = p * r2 - 1.0 / 39916800.0;         // 1/11!
    p = p * r2 + 1.0 / 362880.0;           // 1/9!
    p = p * r2 - 1.0 / 5040.0;
    p = p * r2 + 1.0 / 120.0;
    p = p * r2 - 1.0 / 6.0;
    return r + r * (r2 * p);
}

KPT_HD double cos_poly(double r) {
    double r2 = r * r;
    double p = 1.0 / 20922789888000.0;     // 1/16!
    p = p * r2 - 1.0 / 87178291200.0;      // 1/14!
    p = p * r2 + 1.0 / 479001600.0;        // 1/12!
    p = p * r2 - 1.0 / 3628800.0;          // 1/10!
    p = p * r2 + 1.0 / 40320.0;
    p = p * r2 - 1.0 / 720.0;
    p = p * r2 + 1.0 / 24.0;
    p = p * r2 - 0.5;
    return 1.0 + r2 * p;
}

// Reduces x by multiples of pi/2; returns the remainder and the quadrant.
KPT_HD double reduce_pio2(float x, int* quadrant) {
    double xd = (double)x;
    double k = __builtin_rint(xd * KPT_TWO_OVER_PI);
    double r = (xd - k * KPT_PIO2_HI) - k * KPT_PIO2_LO;
    int64_t ki = (int64_t)k;
    *quadrant = (int)(ki & 3);
    return r;
}

KPT_HD double sin_d(float x) {
    int q; double r = reduce_pio2(x, &q);
    double s = (q & 1) ? cos_poly(r) : sin_poly(r);
    return (q & 2) ? -s : s;
}

KPT_HD double cos_d(float x) {
    int q; double r = reduce_pio2(x, &q);
    double c = (q & 1) ? sin_poly(r) : cos_poly(r);
    return ((q + 1) & 2) ? -c : c;
}

KPT_HD float sinf(float x) {
    if (f_isnan(x) || f_isinf(x)) return f_nan();
    if (__builtin_fabsf(x) > 1.0e6f) return f_nan();  // outside the supported domain
    return (float)sin_d(x);
}

KPT_HD float cosf(float x) {
    if (f_isnan(x) || f_isinf(x)) return f_nan();
    if (__builtin_fabsf(x) > 1.0e6f) return f_nan();
    return (float)cos_d(x);
}

KPT_HD float tanf(float x) {
    if (f_isnan(x) || f_isinf(x)) return f_nan();
    if (__builtin_fabsf(x) > 1.0e6f) return f_nan();
    return (float)(sin_d(x) / cos_d(x));
}

// asin on |x| <= 1/2 by its power series, in double.
KPT_HD double asin_series(double x) {
    double x2 = x * x;
    // coefficients c_n = (2n)! / (4^n (n!)^2 (2n+1)), n = 0..24, folded below
    double c = 1.0;         // running (2n)!/(4^n (n!)^2)
    double term_sum = 0.0;
    double xp = x;          // x^(2n+1)
    for (int n = 0; n < 25; ++n) {
        term_sum = term_sum + c * xp / (double)(2 * n + 1);
        c = c * (double)(2 * n + 1) / (double)(2 * n + 2);
        xp = xp * x2;
    }
    return term_sum;
}

KPT_HD float acosf(float x) {
    if (f_isnan(x)) return x;
    if (x > 1.0f || x < -1.0f) return f_nan();
    const double PIO2 = 1.57079632679489661923;
    double xd = (double)x;
    double ax = xd < 0.0 ? -xd : xd;
    double as;
    if (ax <= 0.5) {
        as = asin_series(ax);
    } else {
        double z = __builtin_sqrt((1.0 - ax) * 0.5);
        as = PIO2 - 2.0 * asin_series(z);
    }
    if (xd < 0.0) as = -as;
    return (float)(PIO2 - as);
}

}  // namespace dm
}  // namespace kpt
