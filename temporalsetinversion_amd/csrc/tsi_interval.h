// tsi_interval.h -- device interval arithmetic for the set-inversion kernels.
//
// Non-outward-rounded binary32 intervals exactly as compute.wgsl:2481-2975 and `fix`
// (3462-3472) define them.  Conventions (the reference leaves them to the driver):
//   * no FMA contraction (the library is built with -ffp-contract=off);
//   * WGSL min/max = IEEE-754-2019 minimumNumber/maximumNumber (__builtin_fminimum_numf):
//     a NaN operand yields the other one, -0 orders below +0;
//   * sqrt and division correctly rounded (hipcc default);
//   * sin/exp/log/atan2/pow: the deterministic routines below, written from their series
//     definitions with only +,-,*,/,floor and exponent-field operations, so every binary32
//     machine computes the same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tsi {

#define TSI_DI __device__ __forceinline__

struct Iv { float lo, hi; };

TSI_DI float wmin(float a, float b) { return __builtin_fminimum_numf(a, b); }
TSI_DI float wmax(float a, float b) { return __builtin_fmaximum_numf(a, b); }
TSI_DI Iv I(float lo, float hi) { return Iv{lo, hi}; }
TSI_DI Iv I_FROM(float a) { return Iv{a, a}; }
TSI_DI bool is_nan_w(float a) { return !(a < 0.0f || a > 0.0f || a == 0.0f); }

constexpr float kBig = 1000000.0f;                 // compute.wgsl:2534
constexpr float kPi = 3.141592653589793f;          // compute.wgsl:2477
constexpr float kTwoPi = 6.283185307179586f;
constexpr float kHalfPi = 1.5707963267948966f;

// ---------------- deterministic transcendental routines ----------------

TSI_DI float pow2i(int k) {
    if (k > 127) return __uint_as_float(0x7f800000u);
    if (k >= -126) return __uint_as_float(uint32_t(k + 127) << 23);
    if (k >= -149) return __uint_as_float(1u << (k + 149));
    return 0.0f;
}

// sin: reduction k = round(x * 2/pi), r = x - k*(P1+P2+P3) (three-part split of pi/2, P1 and P2
// with 11 significant bits), Taylor series of sin r / cos r to degree 9 / 10.
TSI_DI float det_sin(float x) {
    if (!(x - x == 0.0f)) return x - x;
    float k = floorf(x * 0.63661977236758134308f + 0.5f);
    float r = ((x - k * 1.5703125f) - k * 4.837512969970703125e-4f) - k * 7.549790126404332e-08f;
    float kq = k - 4.0f * floorf(k * 0.25f);
    int q = int(kq);
    float r2 = r * r;
    float s = r + r * r2 * (-1.66666666666666666667e-1f + r2 * (8.33333333333333333333e-3f +
              r2 * (-1.98412698412698412698e-4f + r2 * 2.75573192239858906526e-6f)));
    float c = (1.0f - 0.5f * r2) + r2 * r2 * (4.16666666666666666667e-2f + r2 * (-1.38888888888888888889e-3f +
              r2 * (2.48015873015873015873e-5f + r2 * -2.75573192239858906526e-7f)));
    q &= 3;
    float v = (q & 1) ? c : s;
    return (q & 2) ? -v : v;
}

// exp: k = round(x / ln2), r = x - k*(LN2_HI + LN2_LO), Taylor series to degree 8, times 2^k.
TSI_DI float det_exp(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return __uint_as_float(0x7f800000u);
    if (x < -103.972084045410f) return 0.0f;
    float k = floorf(x * 1.44269504088896340736f + 0.5f);
    float r = (x - k * 0.693145751953125f) - k * 1.428606765330187e-06f;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (1.66666666666666666667e-1f + r * (4.16666666666666666667e-2f +
              r * (8.33333333333333333333e-3f + r * (1.38888888888888888889e-3f + r * (1.98412698412698412698e-4f +
              r * 2.48015873015873015873e-5f)))))));
    int ki = int(k);
    if (ki > 127) return (p * pow2i(127)) * pow2i(ki - 127);
    if (ki < -126) return (p * pow2i(ki + 64)) * pow2i(-64);
    return p * pow2i(ki);
}

// log: x = m * 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)) to degree 11.
TSI_DI float det_log(float x) {
    if (x != x) return x;
    if (x < 0.0f) return __uint_as_float(0x7fc00000u);
    if (x == 0.0f) return __uint_as_float(0xff800000u);
    if (x == __uint_as_float(0x7f800000u)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }
    uint32_t u = __float_as_uint(x);
    e += int((u >> 23) & 0xffu) - 127;
    float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356237309504880f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = s2 * (3.33333333333333333333e-1f + s2 * (2.0e-1f + s2 * (1.42857142857142857143e-1f +
              s2 * (1.11111111111111111111e-1f + s2 * 9.09090909090909090909e-2f))));
    float lm = (s + s * p) * 2.0f;
    float fe = float(e);
    return fe * 0.693145751953125f + (fe * 1.428606765330187e-06f + lm);
}

TSI_DI float det_atan01(float t) {
    float off = 0.0f;
    if (t > 0.41421356237309504880f) { t = (t - 1.0f) / (t + 1.0f); off = 0.78539816339744830962f; }
    float t2 = t * t;
    float p = t2 * (-3.33333333333333333333e-1f + t2 * (2.0e-1f + t2 * (-1.42857142857142857143e-1f +
              t2 * (1.11111111111111111111e-1f + t2 * (-9.09090909090909090909e-2f + t2 * (7.69230769230769230769e-2f +
              t2 * (-6.66666666666666666667e-2f + t2 * (5.88235294117647058824e-2f + t2 * (-5.26315789473684210526e-2f +
              t2 * 4.76190476190476190476e-2f)))))))));
    return off + (t + t * p);
}

// atan2(y, x) with C's special cases, octant reduction to atan on [0, 1].
TSI_DI float det_atan2(float y, float x) {
    if (x != x || y != y) return x + y;
    float ax = fabsf(x), ay = fabsf(y);
    bool xneg = (__float_as_uint(x) & 0x80000000u) != 0;
    bool yneg = (__float_as_uint(y) & 0x80000000u) != 0;
    float inf = __uint_as_float(0x7f800000u), a;
    if (ay == 0.0f) a = xneg ? 3.14159265358979323846f : 0.0f;
    else if (ax == 0.0f) a = 1.57079632679489661923f;
    else if (ax == inf && ay == inf) a = xneg ? 3.0f * 0.78539816339744830962f : 0.78539816339744830962f;
    else if (ax == inf) a = xneg ? 3.14159265358979323846f : 0.0f;
    else if (ay == inf) a = 1.57079632679489661923f;
    else {
        if (ay <= ax) a = det_atan01(ay / ax);
        else a = 1.57079632679489661923f - det_atan01(ax / ay);
        if (xneg) a = 3.14159265358979323846f - a;
    }
    return yneg ? -a : a;
}

// x^n by binary exponentiation (POW_II only ever raises to an integer).
TSI_DI float det_powi(float x, int n) {
    uint32_t un = n < 0 ? 0u - uint32_t(n) : uint32_t(n);
    float r = 1.0f, b = x;
    while (un) {
        if (un & 1u) r = r * b;
        un >>= 1;
        if (un) b = b * b;
    }
    return n < 0 ? 1.0f / r : r;
}

// pow(x, 0.45) of get_color_supersample (compute.wgsl:4643-4645) for x in [0, 1].
TSI_DI float det_pow045(float x) {
    if (x != x) return x;
    if (x <= 0.0f) return 0.0f;
    if (x == 1.0f) return 1.0f;
    return det_exp(0.449999988079071045f * det_log(x));
}

// WGSL i32(f32): NaN -> 0, saturating.
TSI_DI int32_t f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return -2147483647 - 1;
    return int32_t(x);
}

// ---------------- interval operations (compute.wgsl line numbers) ----------------

TSI_DI Iv I_FROM_UNORDERED(float a, float b) { return I(wmin(a, b), wmax(a, b)); }       // 2494
TSI_DI bool I_CONTAINS_ZERO(Iv a) { return a.lo <= 0.0f && 0.0f <= a.hi; }              // 2498
TSI_DI float DIAMETER(Iv a) { return a.hi - a.lo; }                                     // 2506
TSI_DI float MID(Iv a) { return (a.lo + a.hi) / 2.0f; }                                  // 2510
TSI_DI Iv ADD_II(Iv a, Iv b) { return I(a.lo + b.lo, a.hi + b.hi); }                   // 2539
TSI_DI Iv ADD_IF(Iv a, float b) { return I(a.lo + b, a.hi + b); }                      // 2543
TSI_DI Iv SUB_II(Iv a, Iv b) { return I(a.lo - b.hi, a.hi - b.lo); }                   // 2547
TSI_DI Iv SUB_IF(Iv a, float b) { return I(a.lo - b, a.hi - b); }                      // 2551
TSI_DI Iv NEG_I(Iv a) { return I(-a.hi, -a.lo); }                                       // 2580

TSI_DI Iv MUL_II(Iv a, Iv b) {                                                           // 2555
    float x = a.lo * b.lo, y = a.lo * b.hi, z = a.hi * b.lo, w = a.hi * b.hi;
    return I(wmin(wmin(x, y), wmin(z, w)), wmax(wmax(x, y), wmax(z, w)));
}
TSI_DI Iv MUL_IF(Iv a, float b) {                                                        // 2563
    return b >= 0.0f ? I(a.lo * b, a.hi * b) : I(a.hi * b, a.lo * b);
}
TSI_DI Iv SQRT_I(Iv a) { return I(sqrtf(wmax(a.lo, 0.0f)), sqrtf(wmax(a.hi, 0.0f))); }  // 2572
TSI_DI Iv EXP_I(Iv a) { return I(det_exp(a.lo), det_exp(a.hi)); }                      // 2576
TSI_DI Iv LOG_I(Iv a) { return I(det_log(wmax(a.lo, 0.0f)), det_log(wmax(a.hi, 0.0f))); }  // 2584
TSI_DI Iv DIV_II(Iv a, Iv b) {                                                           // 2598
    float x = a.lo / b.lo, y = a.lo / b.hi, z = a.hi / b.lo, w = a.hi / b.hi;
    return I(wmin(wmin(x, y), wmin(z, w)), wmax(wmax(x, y), wmax(z, w)));
}
TSI_DI Iv DIV_IF(Iv a, float b) {                                                        // 2609 (= DIV_II(a, [b,b]))
    float x = a.lo / b, z = a.hi / b;
    return I(wmin(wmin(x, x), wmin(z, z)), wmax(wmax(x, x), wmax(z, z)));
}
TSI_DI Iv MIN_II(Iv a, Iv b) { return I(wmin(a.lo, b.lo), wmin(a.hi, b.hi)); }          // 2613
TSI_DI Iv MIN_IF(Iv a, float b) { return I(wmin(a.lo, b), wmin(a.hi, b)); }             // 2617
TSI_DI Iv MAX_IF(Iv a, float b) { return I(wmax(a.lo, b), wmax(a.hi, b)); }             // 2621
TSI_DI Iv MAX_II(Iv a, Iv b) { return I(wmax(a.lo, b.lo), wmax(a.hi, b.hi)); }          // 2625
TSI_DI Iv ABS_I(Iv a) {                                                                  // 2629
    if (I_CONTAINS_ZERO(a)) return I(0.0f, wmax(-a.lo, a.hi));
    if (a.hi < 0.0f) return I(-a.hi, -a.lo);
    return a;
}
TSI_DI Iv ABS_I_RR(Iv a) {                                                               // 2640
    if (I_CONTAINS_ZERO(a)) return I(-1.0f, 1.0f);
    if (a.hi < 0.0f) return I_FROM(-1.0f);
    return I_FROM(1.0f);
}
TSI_DI Iv I_EMPTY() { return I(__uint_as_float(0x7f800000u), __uint_as_float(0xff800000u)); }  // 2655
TSI_DI Iv CAP_II(Iv a, Iv b) {                                                           // 2664
    if (a.hi < b.lo || b.hi < a.lo) return I_EMPTY();
    return I(wmax(a.lo, b.lo), wmin(a.hi, b.hi));
}
TSI_DI Iv SQR_I(Iv a) {                                                                  // 2689
    float x = a.lo * a.lo, y = a.hi * a.hi;
    if (I_CONTAINS_ZERO(a)) return I(0.0f, wmax(x, y));
    return I(wmin(x, y), wmax(x, y));
}
TSI_DI Iv FLOOR_I(Iv a) { return I(floorf(a.lo), floorf(a.hi)); }                       // 2704
TSI_DI Iv FLOOR_I_RR(Iv a) {                                                             // 2708
    if (floorf(a.hi) != floorf(a.lo)) return I(-kBig, kBig);
    return I_FROM(0.0f);
}

TSI_DI Iv ATAN2_II(Iv x, Iv y) {                                        // 2724
    Iv range_atan = I(-kPi, kPi), half = I(-kHalfPi, kHalfPi);
    if (y.lo == 0.0f) return I(0.0f, y.hi);
    if (y.hi == 0.0f) return I(y.lo, 0.0f);
    if (x.lo == 0.0f && x.hi == 0.0f) {
        if (y.lo == 0.0f && y.hi == 0.0f) return I_EMPTY();
        if (y.lo >= 0.0f) return I_FROM(kHalfPi);
        if (y.hi <= 0.0f) return I_FROM(-kHalfPi);
        return half;
    } else if (x.lo > 0.0f) {
        if (y.lo == 0.0f && y.hi == 0.0f) return y;
        if (y.lo >= 0.0f) return I(det_atan2(y.lo, x.hi), det_atan2(y.hi, x.lo));
        if (y.hi <= 0.0f) return I(det_atan2(y.lo, x.lo), det_atan2(y.hi, x.hi));
        return I(det_atan2(y.lo, x.lo), det_atan2(y.hi, x.lo));
    } else if (x.hi < 0.0f) {
        if (y.lo == 0.0f && y.hi == 0.0f) return I_FROM(kPi);
        if (y.lo >= 0.0f) return I(det_atan2(y.hi, x.hi), det_atan2(y.lo, x.lo));
        if (y.hi < 0.0f) return I(det_atan2(y.hi, x.lo), det_atan2(y.lo, x.hi));
        return range_atan;
    } else {
        if (x.lo == 0.0f) {
            if (y.lo == 0.0f && y.hi == 0.0f) return y;
            if (y.lo >= 0.0f) return I(det_atan2(y.lo, x.hi), half.hi);
            if (y.hi <= 0.0f) return I(half.lo, det_atan2(y.hi, x.hi));
            return half;
        } else if (x.hi == 0.0f) {
            if (y.lo == 0.0f && y.hi == 0.0f) return I_FROM(kPi);
            if (y.lo >= 0.0f) return I(kHalfPi, det_atan2(y.lo, x.lo));
            if (y.hi < 0.0f) return I(det_atan2(y.hi, x.lo), -kHalfPi);
            return range_atan;
        } else {
            if (y.lo >= 0.0f) return I(det_atan2(y.lo, x.hi), det_atan2(y.lo, x.lo));
            if (y.hi < 0.0f) return I(det_atan2(y.hi, x.lo), det_atan2(y.hi, x.hi));
            return range_atan;
        }
    }
}

TSI_DI Iv MOD_II(Iv a, Iv b) { return SUB_II(a, MUL_II(b, FLOOR_I(DIV_II(a, b)))); }    // 2818

TSI_DI Iv POW_II(Iv a, Iv b) {                                          // 2828
    int32_t power = f2i_sat(rintf(b.lo));
    float lower, higher;
    if (a.lo <= 0.0f) lower = (power % 2 == 0) ? det_powi(-a.lo, power) : -det_powi(-a.lo, power);
    else lower = det_powi(a.lo, power);
    if (a.hi <= 0.0f) higher = (power % 2 == 0) ? det_powi(-a.hi, power) : -det_powi(-a.hi, power);
    else higher = det_powi(a.hi, power);
    if (I_CONTAINS_ZERO(a)) return I(0.0f, wmax(lower, higher));
    return I(wmin(lower, higher), wmax(lower, higher));
}

TSI_DI float mymod(float x, float y) { return x - y * floorf(x / y); }                  // 2920

TSI_DI Iv SIN_I(Iv a) {                                                 // 2926
    Iv whole = I(-1.0f, 1.0f);
    if (DIAMETER(a) > kTwoPi) return whole;
    float lq = floorf(a.lo / kHalfPi), hq = floorf(a.hi / kHalfPi);
    if (hq - lq > 4.0f) return whole;
    lq = mymod(lq, 4.0f);
    hq = mymod(hq, 4.0f);
    if (lq == hq) {
        if (a.hi - a.lo > kPi) return whole;
        return I_FROM_UNORDERED(det_sin(a.lo), det_sin(a.hi));
    } else if ((lq == 3.0f && hq == 0.0f) || (lq == 1.0f && hq == 2.0f)) {
        return I_FROM_UNORDERED(det_sin(a.lo), det_sin(a.hi));
    } else if ((lq == 0.0f || lq == 3.0f) && (hq == 1.0f || hq == 2.0f)) {
        return I(wmin(det_sin(a.lo), det_sin(a.hi)), 1.0f);
    } else if ((lq == 1.0f || lq == 2.0f) && (hq == 3.0f || hq == 0.0f)) {
        return I(-1.0f, wmax(det_sin(a.lo), det_sin(a.hi)));
    }
    return whole;
}
TSI_DI Iv COS_I(Iv a) { return SIN_I(ADD_II(a, I_FROM(kHalfPi))); }                     // 2959

TSI_DI Iv fix(Iv x) {                                                                    // 3462
    // is_nan -> +-BIG, then clamp; minimumNumber/maximumNumber would drop a NaN the same way,
    // the explicit form keeps the reference's structure.
    float lo = is_nan_w(x.lo) ? -kBig : x.lo;
    float hi = is_nan_w(x.hi) ? kBig : x.hi;
    return I(wmax(lo, -kBig), wmin(hi, kBig));
}

}  // namespace tsi
