// Deterministic elementary functions for the surf-beat step.
//
// The reference evaluates exp/log/pow/sinh/tanh/sin/cos with the host libm
// on its CPU path and with the CUDA device library on its GPU path; the two
// differ in the last bit now and then, and in the chaotic surf-zone flow a
// one-ulp difference grows to ~1e-3 relative after 1000 steps.  To make the
// HIP step reproduce its CPU oracle bit for bit, both use these functions.
//
// Every function is evaluated in double precision with only IEEE add,
// subtract, multiply, divide and (where noted) sqrt, in a fixed operation
// order and without FMA contraction (both sides compile with
// -ffp-contract=off), so host and device produce identical bits.  The float
// entry points round the double result once; they agree with a correctly
// rounded libm except in double-rounding cases (about one input in 2^28).
//
// Methods (published):
//   exp   Cody-Waite reduction by ln2 and the rational form of the remainder
//         used by fdlibm/FreeBSD e_exp.c (Remez polynomial on [0, 0.347])
//   log   reduction to [sqrt(1/2), sqrt(2)), s = f/(2+f), Remez series of
//         fdlibm e_log.c
//   sin/cos  reduction by pi/2 (two-part constant, valid |x| < 2^19), kernel
//         polynomials of fdlibm k_sin.c / k_cos.c on |r| <= pi/4
//   expm1 fdlibm s_expm1.c rational form on |x| <= 1, exp(x)-1 outside
//   sinh  Taylor series on |x| < 1, (e^x - e^-x)/2 otherwise
//   tanh  expm1(2|x|) / (expm1(2|x|) + 2), +-1 beyond |x| = 22
//   pow   exp(y * log|x|) with the integer-exponent sign rule; y = 2 and y = 3
//         as products in double
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define XM_FN __host__ __device__ inline
#else
#define XM_FN inline
#endif

namespace xm {

XM_FN uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
XM_FN double dbl(uint64_t b) { return __builtin_bit_cast(double, b); }
XM_FN bool isnan_(double x) { return x != x; }
XM_FN double inf_() { return dbl(0x7ff0000000000000ull); }
XM_FN double nan_() { return dbl(0x7ff8000000000000ull); }

// x * 2^n for the exponent range of exp (|n| <= 1100), two exact steps.
XM_FN double scale2(double x, int n) {
    if (n > 1023) {
        x = x * dbl(0x7fe0000000000000ull);  // 2^1023
        n -= 1023;
        if (n > 1023) n = 1023;
    } else if (n < -1022) {
        x = x * dbl(0x0010000000000000ull);  // 2^-1022
        n += 1022;
        if (n < -1022) n = -1022;
    }
    return x * dbl((uint64_t)(n + 1023) << 52);
}

XM_FN int round_int(double v) { return (int)(v >= 0.0 ? v + 0.5 : v - 0.5); }

XM_FN double exp(double x) {
    if (isnan_(x)) return x;
    if (x > 709.782712893383973096) return inf_();
    if (x < -745.133219101941108420) return 0.0;
    const double ln2hi = 6.93147180369123816490e-01, ln2lo = 1.90821492927058770002e-10;
    const double invln2 = 1.44269504088896338700e+00;
    int k = round_int(x * invln2);
    double hi = x - k * ln2hi;  // exact: ln2hi has 32 significant bits
    double lo = k * ln2lo;
    double r = hi - lo;
    const double P1 = 1.66666666666666019037e-01, P2 = -2.77777777770155933842e-03,
                 P3 = 6.61375632143793436117e-05, P4 = -1.65339022054652515390e-06,
                 P5 = 4.13813679705723846039e-08;
    double t = r * r;
    double c = r - t * (P1 + t * (P2 + t * (P3 + t * (P4 + t * P5))));
    double y = 1.0 - ((lo - (r * c) / (2.0 - c)) - hi);
    return scale2(y, k);
}

XM_FN double log(double x) {
    if (isnan_(x)) return x;
    if (x < 0.0) return nan_();
    if (x == 0.0) return -inf_();
    if (x == inf_()) return x;
    uint64_t b = bits(x);
    int k = 0;
    if ((b >> 52) == 0) {  // subnormal
        x = x * dbl(0x4350000000000000ull);  // 2^54
        b = bits(x);
        k = -54;
    }
    k += (int)(b >> 52) - 1023;
    double m = dbl((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1,2)
    if (m > 1.41421356237309504880) {
        m = m * 0.5;
        k += 1;
    }
    double f = m - 1.0;
    const double ln2hi = 6.93147180369123816490e-01, ln2lo = 1.90821492927058770002e-10;
    const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01,
                 Lg3 = 2.857142874366239149e-01, Lg4 = 2.222219843214978396e-01,
                 Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                 Lg7 = 1.479819860511658591e-01;
    double s = f / (2.0 + f);
    double z = s * s;
    double w = z * z;
    double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
    double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
    double R = t2 + t1;
    double hfsq = 0.5 * f * f;
    double dk = (double)k;
    if (k == 0) return f - (hfsq - s * (hfsq + R));
    return dk * ln2hi - ((hfsq - (s * (hfsq + R) + dk * ln2lo)) - f);
}

XM_FN double log10(double x) { return log(x) / 2.30258509299404568402; }

XM_FN double expm1(double x) {
    if (isnan_(x)) return x;
    if (x > 1.0 || x < -1.0) return exp(x) - 1.0;
    // fdlibm s_expm1.c on |x| <= 1: reduction by one ln2 above 0.5 ln2,
    // rational approximation of the remainder, exact reconstruction
    const double ln2hi = 6.93147180369123816490e-01, ln2lo = 1.90821492927058770002e-10;
    const double Q1 = -3.33333333333331316428e-02, Q2 = 1.58730158725481460165e-03,
                 Q3 = -7.93650757867487942473e-05, Q4 = 4.00821782732936239552e-06,
                 Q5 = -2.01099218183624371326e-07;
    double ax = x < 0.0 ? -x : x;
    int k = 0;
    double c = 0.0;
    if (ax > 0.5 * 6.93147180559945286227e-01) {
        double hi, lo;
        if (x > 0.0) {
            hi = x - ln2hi;
            lo = ln2lo;
            k = 1;
        } else {
            hi = x + ln2hi;
            lo = -ln2lo;
            k = -1;
        }
        x = hi - lo;
        c = (hi - x) - lo;
    } else if (ax < 5.55111512312578270212e-17) {  // 2^-54
        return x;
    }
    double hfx = 0.5 * x;
    double hxs = x * hfx;
    double r1 = 1.0 + hxs * (Q1 + hxs * (Q2 + hxs * (Q3 + hxs * (Q4 + hxs * Q5))));
    double t = 3.0 - r1 * hfx;
    double e = hxs * ((r1 - t) / (6.0 - x * t));
    if (k == 0) return x - (x * e - hxs);
    e = (x * (e - c) - c);
    e = e - hxs;
    if (k == -1) return 0.5 * (x - e) - 0.5;
    if (x < -0.25) return -2.0 * (e - (x + 0.5));
    return 1.0 + 2.0 * (x - e);
}

XM_FN double sinh(double x) {
    if (isnan_(x)) return x;
    double a = x < 0.0 ? -x : x;
    if (a < 1.0) {
        // x * sum_{k=0..11} z^k/(2k+1)!, z = x^2, coefficients rounded to double
        const double c[12] = {1.00000000000000000e+00, 1.66666666666666657e-01, 8.33333333333333322e-03, 1.98412698412698413e-04, 2.75573192239858925e-06, 2.50521083854417202e-08, 1.60590438368216133e-10, 7.64716373181981641e-13, 2.81145725434552060e-15, 8.22063524662432950e-18, 1.95729410633912626e-20, 3.86817017063068413e-23};
        double z = x * x;
        double s = c[11];
        for (int k = 10; k >= 0; k--) s = c[k] + s * z;
        return x * s;
    }
    if (a > 709.0) {
        double h = exp(0.5 * a);
        double r = 0.5 * h * h;
        return x < 0.0 ? -r : r;
    }
    double e = exp(a);
    double r = 0.5 * (e - 1.0 / e);
    return x < 0.0 ? -r : r;
}

XM_FN double tanh(double x) {
    if (isnan_(x)) return x;
    double a = x < 0.0 ? -x : x;
    double r;
    if (a >= 22.0) {
        r = 1.0;
    } else {
        double t = expm1(2.0 * a);
        r = t / (t + 2.0);
    }
    return x < 0.0 ? -r : r;
}

XM_FN double ksin(double r) {
    const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                 S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                 S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    double z = r * r;
    double v = z * r;
    double q = S2 + z * (S3 + z * (S4 + z * (S5 + z * S6)));
    return r + v * (S1 + z * q);
}
XM_FN double kcos(double r) {
    const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                 C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                 C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    double z = r * r;
    double q = z * (C1 + z * (C2 + z * (C3 + z * (C4 + z * (C5 + z * C6)))));
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    return w + (((1.0 - w) - hz) + z * q);
}
// reduce by pi/2: r in [-pi/4, pi/4], quadrant q
XM_FN double rpio2(double x, int& q) {
    const double invpio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00, pio2_1t = 6.07710050650619224932e-11;
    int k = round_int(x * invpio2);
    double r = (x - k * pio2_1) - k * pio2_1t;
    q = k & 3;
    return r;
}
XM_FN double sin(double x) {
    if (isnan_(x)) return x;
    int q;
    double r = rpio2(x, q);
    switch (q) {
        case 0: return ksin(r);
        case 1: return kcos(r);
        case 2: return -ksin(r);
        default: return -kcos(r);
    }
}
XM_FN double cos(double x) {
    if (isnan_(x)) return x;
    int q;
    double r = rpio2(x, q);
    switch (q) {
        case 0: return kcos(r);
        case 1: return -ksin(r);
        case 2: return -kcos(r);
        default: return ksin(r);
    }
}

XM_FN bool is_int(double y) {
    if (y != y || y == inf_() || y == -inf_()) return false;
    double a = y < 0 ? -y : y;
    if (a >= 4503599627370496.0) return true;  // 2^52
    double t = (double)(int64_t)a;
    return t == a;
}
XM_FN bool is_odd_int(double y) {
    if (!is_int(y)) return false;
    double a = y < 0 ? -y : y;
    if (a >= 9007199254740992.0) return false;  // 2^53: even
    return ((int64_t)a & 1) != 0;
}

XM_FN double pow(double x, double y) {
    if (y == 0.0) return 1.0;
    if (x == 1.0) return 1.0;
    if (isnan_(x) || isnan_(y)) return nan_();
    if (y == 2.0) return x * x;
    if (y == 3.0) return x * x * x;
    if (x == 0.0) {
        if (y > 0.0) return is_odd_int(y) ? x : 0.0;
        return is_odd_int(y) && bits(x) >> 63 ? -inf_() : inf_();
    }
    double a = x;
    double sign = 1.0;
    if (x < 0.0) {
        if (!is_int(y)) return nan_();
        a = -x;
        if (is_odd_int(y)) sign = -1.0;
    }
    if (a == inf_()) return y > 0.0 ? sign * inf_() : sign * 0.0;
    return sign * exp(y * log(a));
}

// ---- float entry points: one rounding of the double result ----
XM_FN float expf(float x) { return (float)exp((double)x); }
XM_FN float logf(float x) { return (float)log((double)x); }
XM_FN float log10f(float x) { return (float)log10((double)x); }
XM_FN float powf(float x, float y) { return (float)pow((double)x, (double)y); }
XM_FN float sinhf(float x) { return (float)sinh((double)x); }
XM_FN float tanhf(float x) { return (float)tanh((double)x); }
XM_FN float sinf(float x) { return (float)sin((double)x); }
XM_FN float cosf(float x) { return (float)cos((double)x); }

}  // namespace xm
