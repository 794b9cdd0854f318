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
//
// Everything here is branch-free: the reference's if/else chains become selects over values
// computed unconditionally.  The tape interpreter dispatches on a wave-uniform opcode; keeping
// the per-lane code free of divergent branches lets the compiler emit that dispatch as plain
// scalar branches instead of exec-mask flow blocks.
#pragma once
#ifdef __HIPCC_RTC__   // hiprtc provides the runtime declarations; fixed-width types live here
using __hip_internal::int32_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

namespace tsi {

#define TSI_DI __device__ __forceinline__

struct Iv { float lo, hi; };

TSI_DI float wmin(float a, float b) { return __builtin_fminimum_numf(a, b); }
TSI_DI float wmax(float a, float b) { return __builtin_fmaximum_numf(a, b); }
TSI_DI float wmin4(float x, float y, float z, float w) { return wmin(wmin(x, y), wmin(z, w)); }
TSI_DI float wmax4(float x, float y, float z, float w) { return wmax(wmax(x, y), wmax(z, w)); }
TSI_DI Iv I(float lo, float hi) { return Iv{lo, hi}; }
TSI_DI Iv I_FROM(float a) { return Iv{a, a}; }
TSI_DI bool is_nan_w(float a) { return !(a < 0.0f || a > 0.0f || a == 0.0f); }
TSI_DI Iv sel(bool c, Iv a, Iv b) { return Iv{c ? a.lo : b.lo, c ? a.hi : b.hi}; }

constexpr float kBig = 1000000.0f;                 // compute.wgsl:2534
constexpr float kPi = 3.141592653589793f;          // compute.wgsl:2477
constexpr float kTwoPi = 6.283185307179586f;
constexpr float kHalfPi = 1.5707963267948966f;

// ---------------- deterministic transcendental routines ----------------

TSI_DI float f_inf() { return __uint_as_float(0x7f800000u); }

// 2^k as a float for k in [-149, 128] (128 -> +inf), built from the exponent field.
TSI_DI float pow2i(int k) {
    const float normal = __uint_as_float(uint32_t(k + 127) << 23);
    const float sub = __uint_as_float(1u << ((k + 149) & 31));
    float r = k >= -126 ? normal : (k >= -149 ? sub : 0.0f);
    return k > 127 ? f_inf() : r;
}

// sin: reduction k = round(x * 2/pi), r = x - k*(P1+P2+P3) (three-part split of pi/2, P1 and P2
// with 11 significant bits), Taylor series of sin r / cos r to degree 9 / 10.
TSI_DI float det_sin(float x) {
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
    v = (q & 2) ? -v : v;
    return (x - x == 0.0f) ? v : x - x;            // NaN, +-inf -> NaN
}

// exp: k = round(x / ln2), r = x - k*(LN2_HI + LN2_LO), Taylor series to degree 8, times 2^k.
TSI_DI float det_exp(float x) {
    const float xc = wmin(wmax(x, -104.0f), 89.0f);   // keeps k in int range; results selected below
    float k = floorf(xc * 1.44269504088896340736f + 0.5f);
    float r = (xc - k * 0.693145751953125f) - k * 1.428606765330187e-06f;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (1.66666666666666666667e-1f + r * (4.16666666666666666667e-2f +
              r * (8.33333333333333333333e-3f + r * (1.38888888888888888889e-3f + r * (1.98412698412698412698e-4f +
              r * 2.48015873015873015873e-5f)))))));
    int ki = int(k);
    const float hiv = (p * pow2i(127)) * pow2i(ki - 127);
    const float lov = (p * pow2i(ki + 64)) * pow2i(-64);
    const float mid = p * pow2i(ki);
    float v = ki > 127 ? hiv : (ki < -126 ? lov : mid);
    v = x < -103.972084045410f ? 0.0f : v;
    v = x > 88.72283935546875f ? f_inf() : v;
    return x != x ? x : v;
}

// log: x = m * 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)) to degree 11.
TSI_DI float det_log(float x) {
    const bool sub = x < 1.17549435e-38f;
    const float xs = sub ? x * 8388608.0f : x;
    int e = sub ? -23 : 0;
    uint32_t u = __float_as_uint(xs);
    e += int((u >> 23) & 0xffu) - 127;
    float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
    const bool big = m > 1.41421356237309504880f;
    m = big ? m * 0.5f : m;
    e += big ? 1 : 0;
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = s2 * (3.33333333333333333333e-1f + s2 * (2.0e-1f + s2 * (1.42857142857142857143e-1f +
              s2 * (1.11111111111111111111e-1f + s2 * 9.09090909090909090909e-2f))));
    float lm = (s + s * p) * 2.0f;
    float fe = float(e);
    float v = fe * 0.693145751953125f + (fe * 1.428606765330187e-06f + lm);
    v = x == f_inf() ? x : v;
    v = x == 0.0f ? __uint_as_float(0xff800000u) : v;
    v = x < 0.0f ? __uint_as_float(0x7fc00000u) : v;
    return x != x ? x : v;
}

TSI_DI float det_atan01(float t) {
    const bool big = t > 0.41421356237309504880f;
    t = big ? (t - 1.0f) / (t + 1.0f) : t;
    const float off = big ? 0.78539816339744830962f : 0.0f;
    float t2 = t * t;
    float p = t2 * (-3.33333333333333333333e-1f + t2 * (2.0e-1f + t2 * (-1.42857142857142857143e-1f +
              t2 * (1.11111111111111111111e-1f + t2 * (-9.09090909090909090909e-2f + t2 * (7.69230769230769230769e-2f +
              t2 * (-6.66666666666666666667e-2f + t2 * (5.88235294117647058824e-2f + t2 * (-5.26315789473684210526e-2f +
              t2 * 4.76190476190476190476e-2f)))))))));
    return off + (t + t * p);
}

// atan2(y, x) with C's special cases, octant reduction to atan on [0, 1].
TSI_DI float det_atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const bool xneg = (__float_as_uint(x) & 0x80000000u) != 0;
    const bool yneg = (__float_as_uint(y) & 0x80000000u) != 0;
    const float inf = f_inf();
    const bool le = ay <= ax;
    const float q = le ? ay / ax : ax / ay;
    const float a0 = det_atan01(q);
    float a = le ? a0 : 1.57079632679489661923f - a0;
    a = xneg ? 3.14159265358979323846f - a : a;
    a = (ay == inf) ? 1.57079632679489661923f : a;
    a = (ax == inf) ? (xneg ? 3.14159265358979323846f : 0.0f) : a;
    a = (ax == inf && ay == inf) ? (xneg ? 3.0f * 0.78539816339744830962f : 0.78539816339744830962f) : a;
    a = (ax == 0.0f) ? 1.57079632679489661923f : a;
    a = (ay == 0.0f) ? (xneg ? 3.14159265358979323846f : 0.0f) : a;
    const float v = yneg ? -a : a;
    return (x != x || y != y) ? x + y : v;
}

// x^n by binary exponentiation (POW_II only ever raises to an integer); a fixed 32-step
// loop so the trip count does not depend on the lane.
TSI_DI float det_powi(float x, int n) {
    uint32_t un = n < 0 ? 0u - uint32_t(n) : uint32_t(n);
    float r = 1.0f, b = x;
#pragma unroll 4
    for (int i = 0; i < 32; ++i) {
        r = (un & 1u) ? r * b : r;
        un >>= 1;
        b = un ? b * b : b;
    }
    return n < 0 ? 1.0f / r : r;
}

// cos: det_sin's reduction with the quadrant advanced by one (cos r = sin(r + pi/2) without
// rounding x + pi/2).
TSI_DI float det_cos(float x) {
    float k = floorf(x * 0.63661977236758134308f + 0.5f);
    float r = ((x - k * 1.5703125f) - k * 4.837512969970703125e-4f) - k * 7.549790126404332e-08f;
    float kq = k - 4.0f * floorf(k * 0.25f);
    int q = int(kq) + 1;
    float r2 = r * r;
    float s = r + r * r2 * (-1.66666666666666666667e-1f + r2 * (8.33333333333333333333e-3f +
              r2 * (-1.98412698412698412698e-4f + r2 * 2.75573192239858906526e-6f)));
    float c = (1.0f - 0.5f * r2) + r2 * r2 * (4.16666666666666666667e-2f + r2 * (-1.38888888888888888889e-3f +
              r2 * (2.48015873015873015873e-5f + r2 * -2.75573192239858906526e-7f)));
    q &= 3;
    float v = (q & 1) ? c : s;
    v = (q & 2) ? -v : v;
    return (x - x == 0.0f) ? v : x - x;
}

// x^y for the shading formulas: exp(y log x) for x > 0; 0^y is 0 (y > 0), 1 (y = 0) or +inf;
// x < 0 gives NaN (GLSL/WGSL pow is undefined there); 1^y = 1.
TSI_DI float det_pow(float x, float y) {
    float v = det_exp(y * det_log(x));
    const float z = y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : f_inf());
    v = x == 0.0f ? z : v;
    v = x == 1.0f ? 1.0f : v;
    v = x < 0.0f ? __uint_as_float(0x7fc00000u) : v;
    return (x != x || y != y) ? x + y : v;
}

// acos x = atan2(sqrt((1 - x)(1 + x)), x) on [-1, 1]; NaN outside.
TSI_DI float det_acos(float x) {
    const float v = det_atan2(sqrtf((1.0f - x) * (1.0f + x)), x);
    return (x > 1.0f || x < -1.0f) ? __uint_as_float(0x7fc00000u) : v;
}

// pow(x, 0.45) of get_color_supersample (compute.wgsl:4643-4645) for x in [0, 1].
TSI_DI float det_pow045(float x) {
    float v = det_exp(0.449999988079071045f * det_log(x));
    v = x == 1.0f ? 1.0f : v;
    v = x <= 0.0f ? 0.0f : v;
    return x != x ? x : v;
}

// WGSL i32(f32): NaN -> 0, saturating.
TSI_DI int32_t f2i_sat(float x) {
    int32_t v = int32_t(wmin(wmax(x, -2147483648.0f), 2147483520.0f));
    v = x >= 2147483648.0f ? 2147483647 : v;
    return x != x ? 0 : v;
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
    return I(wmin4(x, y, z, w), wmax4(x, y, z, w));
}
TSI_DI Iv MUL_IF(Iv a, float b) {                                                        // 2563
    const float x = a.lo * b, y = a.hi * b;
    return b >= 0.0f ? I(x, y) : I(y, x);
}
// Correctly rounded square root of x in [-0, +inf] (SQRT_I's operands: wmax(v, 0) maps NaN and
// negative values to +0).  The method of LLVM's AMDGPU lowering of a correctly rounded binary32
// sqrt: inputs below 2^-96 scaled by 2^32 (exact), the hardware square root s (within 1 ulp), the
// neighbour s - 1 ulp taken when x - (s - 1 ulp) s <= 0 and s + 1 ulp when x - (s + 1 ulp) s > 0
// (exact residuals by fma), the result scaled back by 2^-16.  The lowering's last step, which
// returns x itself for +-0 and +inf, is left out: for x = +-0 and +inf both residual tests fail
// (a NaN neighbour or residual, or a zero residual) and s, which is then x, is returned.  Equal,
// bit for bit, to sqrtf for every binary32 in [-0, +inf] and every NaN (an exhaustive GPU test).
TSI_DI float sqrt_nonneg(float x) {
    const bool tiny = x < 0x1p-96f;
    const float xs = tiny ? x * 0x1p32f : x;
    const float s = __builtin_amdgcn_sqrtf(xs);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u);
    const float su = __uint_as_float(__float_as_uint(s) + 1u);
    float r = __builtin_fmaf(-sd, s, xs) <= 0.0f ? sd : s;
    r = __builtin_fmaf(-su, s, xs) > 0.0f ? su : r;
    return tiny ? r * 0x1p-16f : r;
}
TSI_DI Iv SQRT_I(Iv a) { return I(sqrt_nonneg(wmax(a.lo, 0.0f)), sqrt_nonneg(wmax(a.hi, 0.0f))); }  // 2572
// SQRT_I of an interval whose bounds are each +0, positive or +inf (never -0, negative or NaN),
// where wmax(v, 0) is v: the compiled modules' form for operands they prove so
TSI_DI Iv SQRT_NN(Iv a) { return I(sqrt_nonneg(a.lo), sqrt_nonneg(a.hi)); }
TSI_DI Iv EXP_I(Iv a) { return I(det_exp(a.lo), det_exp(a.hi)); }                      // 2576
TSI_DI Iv LOG_I(Iv a) { return I(det_log(wmax(a.lo, 0.0f)), det_log(wmax(a.hi, 0.0f))); }  // 2584
TSI_DI Iv DIV_II(Iv a, Iv b) {                                                           // 2598
    float x = a.lo / b.lo, y = a.lo / b.hi, z = a.hi / b.lo, w = a.hi / b.hi;
    return I(wmin4(x, y, z, w), wmax4(x, y, z, w));
}
TSI_DI Iv DIV_IF(Iv a, float b) {                                                        // 2609 (= DIV_II(a, [b,b]))
    float x = a.lo / b, z = a.hi / b;
    return I(wmin(x, z), wmax(x, z));
}
TSI_DI Iv MIN_II(Iv a, Iv b) { return I(wmin(a.lo, b.lo), wmin(a.hi, b.hi)); }          // 2613
TSI_DI Iv MIN_IF(Iv a, float b) { return I(wmin(a.lo, b), wmin(a.hi, b)); }             // 2617
TSI_DI Iv MAX_IF(Iv a, float b) { return I(wmax(a.lo, b), wmax(a.hi, b)); }             // 2621
TSI_DI Iv MAX_II(Iv a, Iv b) { return I(wmax(a.lo, b.lo), wmax(a.hi, b.hi)); }          // 2625
TSI_DI Iv ABS_I(Iv a) {                                                                  // 2629
    const Iv z = I(0.0f, wmax(-a.lo, a.hi)), n = I(-a.hi, -a.lo);
    return sel(I_CONTAINS_ZERO(a), z, sel(a.hi < 0.0f, n, a));
}
TSI_DI Iv ABS_I_RR(Iv a) {                                                               // 2640
    return sel(I_CONTAINS_ZERO(a), I(-1.0f, 1.0f), sel(a.hi < 0.0f, I_FROM(-1.0f), I_FROM(1.0f)));
}
TSI_DI Iv I_EMPTY() { return I(f_inf(), __uint_as_float(0xff800000u)); }                // 2655
TSI_DI Iv CAP_II(Iv a, Iv b) {                                                           // 2664
    return sel(a.hi < b.lo || b.hi < a.lo, I_EMPTY(), I(wmax(a.lo, b.lo), wmin(a.hi, b.hi)));
}
TSI_DI Iv SQR_I(Iv a) {                                                                  // 2689
    float x = a.lo * a.lo, y = a.hi * a.hi;
    const float mx = wmax(x, y);
    return I(I_CONTAINS_ZERO(a) ? 0.0f : wmin(x, y), mx);
}
TSI_DI Iv FLOOR_I(Iv a) { return I(floorf(a.lo), floorf(a.hi)); }                       // 2704
TSI_DI Iv FLOOR_I_RR(Iv a) {                                                             // 2708
    return sel(floorf(a.hi) != floorf(a.lo), I(-kBig, kBig), I_FROM(0.0f));
}

TSI_DI Iv ATAN2_II(Iv x, Iv y) {                                                         // 2724
    // the four corner angles the branches of the reference choose from
    const float A = det_atan2(y.lo, x.lo), B = det_atan2(y.lo, x.hi);
    const float Cc = det_atan2(y.hi, x.lo), D = det_atan2(y.hi, x.hi);
    const Iv range = I(-kPi, kPi), half = I(-kHalfPi, kHalfPi);
    const bool y0 = y.lo == 0.0f && y.hi == 0.0f;
    const Iv r_x0 = sel(y0, I_EMPTY(), sel(y.lo >= 0.0f, I_FROM(kHalfPi), sel(y.hi <= 0.0f, I_FROM(-kHalfPi), half)));
    const Iv r_xpos = sel(y0, y, sel(y.lo >= 0.0f, I(B, Cc), sel(y.hi <= 0.0f, I(A, D), I(A, Cc))));
    const Iv r_xneg = sel(y0, I_FROM(kPi), sel(y.lo >= 0.0f, I(D, A), sel(y.hi < 0.0f, I(Cc, B), range)));
    const Iv r_xlo0 = sel(y0, y, sel(y.lo >= 0.0f, I(B, half.hi), sel(y.hi <= 0.0f, I(half.lo, D), half)));
    const Iv r_xhi0 = sel(y0, I_FROM(kPi), sel(y.lo >= 0.0f, I(kHalfPi, A), sel(y.hi < 0.0f, I(Cc, -kHalfPi), range)));
    const Iv r_else = sel(y.lo >= 0.0f, I(B, A), sel(y.hi < 0.0f, I(Cc, D), range));
    const Iv r_mixed = sel(x.lo == 0.0f, r_xlo0, sel(x.hi == 0.0f, r_xhi0, r_else));
    Iv r = sel(x.lo == 0.0f && x.hi == 0.0f, r_x0, sel(x.lo > 0.0f, r_xpos, sel(x.hi < 0.0f, r_xneg, r_mixed)));
    r = sel(y.hi == 0.0f, I(y.lo, 0.0f), r);
    return sel(y.lo == 0.0f, I(0.0f, y.hi), r);
}

TSI_DI Iv MOD_II(Iv a, Iv b) { return SUB_II(a, MUL_II(b, FLOOR_I(DIV_II(a, b)))); }    // 2818

TSI_DI Iv POW_II(Iv a, Iv b) {                                                           // 2828
    int32_t power = f2i_sat(rintf(b.lo));
    const bool even = power % 2 == 0;
    const float pl = det_powi(a.lo <= 0.0f ? -a.lo : a.lo, power);
    const float ph = det_powi(a.hi <= 0.0f ? -a.hi : a.hi, power);
    const float lower = a.lo <= 0.0f ? (even ? pl : -pl) : pl;
    const float higher = a.hi <= 0.0f ? (even ? ph : -ph) : ph;
    return sel(I_CONTAINS_ZERO(a), I(0.0f, wmax(lower, higher)), I(wmin(lower, higher), wmax(lower, higher)));
}

TSI_DI float mymod(float x, float y) { return x - y * floorf(x / y); }                  // 2920

TSI_DI Iv SIN_I(Iv a) {                                                                  // 2926
    const Iv whole = I(-1.0f, 1.0f);
    float lq = floorf(a.lo / kHalfPi), hq = floorf(a.hi / kHalfPi);
    const bool too_wide = DIAMETER(a) > kTwoPi || hq - lq > 4.0f;
    lq = mymod(lq, 4.0f);
    hq = mymod(hq, 4.0f);
    const float sl = det_sin(a.lo), sh = det_sin(a.hi);
    const Iv unord = I_FROM_UNORDERED(sl, sh);
    const bool same = lq == hq;
    const bool c2 = (lq == 3.0f && hq == 0.0f) || (lq == 1.0f && hq == 2.0f);
    const bool c3 = (lq == 0.0f || lq == 3.0f) && (hq == 1.0f || hq == 2.0f);
    const bool c4 = (lq == 1.0f || lq == 2.0f) && (hq == 3.0f || hq == 0.0f);
    Iv r = sel(c4, I(-1.0f, wmax(sl, sh)), whole);
    r = sel(c3, I(wmin(sl, sh), 1.0f), r);
    r = sel(c2, unord, r);
    r = sel(same, sel(a.hi - a.lo > kPi, whole, unord), r);
    return sel(too_wide, whole, r);
}
TSI_DI Iv COS_I(Iv a) { return SIN_I(ADD_II(a, I_FROM(kHalfPi))); }                     // 2959

// compute.wgsl:3462-3472: a NaN lo becomes -BIG and a NaN hi +BIG, then lo is clamped from below
// and hi from above.  maximumNumber / minimumNumber return the non-NaN operand, so the clamp
// alone performs the NaN replacement: two operations per interval instead of six.
TSI_DI Iv fix(Iv x) { return I(wmax(x.lo, -kBig), wmin(x.hi, kBig)); }

}  // namespace tsi
