/*
 * detmath.c -- deterministic f32 transcendental routines of the oracle.
 * TEST INFRASTRUCTURE ONLY (see tsi_oracle.h).
 *
 * The reference calls WGSL sin/exp/log/pow/atan2 (compute.wgsl:2572-2970, 4643-4645),
 * whose results depend on the Vulkan driver.  The oracle and the HIP kernels both use
 * the routines below, written from the textbook definitions:
 *   sin  : Cody-Waite reduction by pi/2 (three-part split) + Taylor series of sin/cos
 *          on [-pi/4, pi/4] (degrees 9 and 10);
 *   exp  : reduction by ln2 (two-part split) + Taylor series to degree 8, scaling by 2^k
 *          built from the exponent field;
 *   log  : exponent extraction, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1))
 *          series to degree 11;
 *   atan2: octant reduction, atan t = pi/4 + atan((t-1)/(t+1)) above tan(pi/8), Taylor
 *          series to degree 21;
 *   powi : binary exponentiation (POW_II only ever takes an integer exponent);
 *   pow045: exp(0.45 * log x), the gamma of rt_main;
 *   cos  : sin's reduction with the quadrant advanced by one;
 *   pow  : exp(y log x) with pow's special cases at 0 and 1;
 *   acos : atan2(sqrt((1 - x)(1 + x)), x).
 * Only +, -, *, /, floor, comparisons and exponent-field bit operations are used, in a
 * fixed order with no contraction, so the result is the same on any IEEE binary32
 * machine.  Accuracy is pinned against numpy float64 in tests/test_oracle_math.py.
 *
 * WGSL min/max: IEEE-754-2019 minimumNumber / maximumNumber.
 */
#include "tsi_oracle.h"
#include <string.h>
#include <math.h>

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

float orc_wmin(float a, float b)
{
    if (a != a) return b;
    if (b != b) return a;
    if (a < b) return a;
    if (b < a) return b;
    return (f2u(a) & 0x80000000u) ? a : b;   /* equal: -0 below +0 */
}

float orc_wmax(float a, float b)
{
    if (a != a) return b;
    if (b != b) return a;
    if (a > b) return a;
    if (b > a) return b;
    return (f2u(a) & 0x80000000u) ? b : a;
}

/* 2^k as a float, k in [-149, 128] (128 -> inf) */
static float pow2i(int k)
{
    if (k > 127) return u2f(0x7f800000u);
    if (k >= -126) return u2f((uint32_t)(k + 127) << 23);
    if (k >= -149) return u2f(1u << (k + 149));
    return 0.0f;
}

#define PIO2_1 1.5703125f
#define PIO2_2 4.837512969970703125e-4f
#define PIO2_3 7.549790126404332e-08f
#define TWO_OVER_PI 0.63661977236758134308f

float orc_det_sinf(float x)
{
    if (!(x - x == 0.0f)) return x - x;          /* NaN, +-inf -> NaN */
    float k = floorf(x * TWO_OVER_PI + 0.5f);
    float r = ((x - k * PIO2_1) - k * PIO2_2) - k * PIO2_3;
    float kq = k - 4.0f * floorf(k * 0.25f);      /* k mod 4, exact for |k| < 2^24 */
    int q = (int)kq;
    float r2 = r * r;
    float s = r + r * r2 * (-1.66666666666666666667e-1f + r2 * (8.33333333333333333333e-3f +
              r2 * (-1.98412698412698412698e-4f + r2 * 2.75573192239858906526e-6f)));
    float c = (1.0f - 0.5f * r2) + r2 * r2 * (4.16666666666666666667e-2f + r2 * (-1.38888888888888888889e-3f +
              r2 * (2.48015873015873015873e-5f + r2 * -2.75573192239858906526e-7f)));
    switch (q & 3) {
    case 0: return s;
    case 1: return c;
    case 2: return -s;
    default: return -c;
    }
}

/* cos: the sine's reduction with the quadrant advanced by one (no rounding of x + pi/2). */
float orc_det_cosf(float x)
{
    if (!(x - x == 0.0f)) return x - x;
    float k = floorf(x * TWO_OVER_PI + 0.5f);
    float r = ((x - k * PIO2_1) - k * PIO2_2) - k * PIO2_3;
    float kq = k - 4.0f * floorf(k * 0.25f);
    int q = (int)kq + 1;
    float r2 = r * r;
    float s = r + r * r2 * (-1.66666666666666666667e-1f + r2 * (8.33333333333333333333e-3f +
              r2 * (-1.98412698412698412698e-4f + r2 * 2.75573192239858906526e-6f)));
    float c = (1.0f - 0.5f * r2) + r2 * r2 * (4.16666666666666666667e-2f + r2 * (-1.38888888888888888889e-3f +
              r2 * (2.48015873015873015873e-5f + r2 * -2.75573192239858906526e-7f)));
    switch (q & 3) {
    case 0: return s;
    case 1: return c;
    case 2: return -s;
    default: return -c;
    }
}

#define LN2_HI 0.693145751953125f
#define LN2_LO 1.428606765330187e-06f
#define LOG2E 1.44269504088896340736f

float orc_det_expf(float x)
{
    if (x != x) return x;
    if (x > 88.72283935546875f) return u2f(0x7f800000u);
    if (x < -103.972084045410f) return 0.0f;
    float k = floorf(x * LOG2E + 0.5f);
    float r = (x - k * LN2_HI) - k * LN2_LO;
    float p = 1.0f + r * (1.0f + r * (0.5f + r * (1.66666666666666666667e-1f + r * (4.16666666666666666667e-2f +
              r * (8.33333333333333333333e-3f + r * (1.38888888888888888889e-3f + r * (1.98412698412698412698e-4f +
              r * 2.48015873015873015873e-5f)))))));
    int ki = (int)k;
    if (ki > 127) return (p * pow2i(127)) * pow2i(ki - 127);
    if (ki < -126) return (p * pow2i(ki + 64)) * pow2i(-64);
    return p * pow2i(ki);
}

#define SQRT2F 1.41421356237309504880f

float orc_det_logf(float x)
{
    if (x != x) return x;
    if (x < 0.0f) return u2f(0x7fc00000u);
    if (x == 0.0f) return u2f(0xff800000u);
    if (x == u2f(0x7f800000u)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 8388608.0f; e = -23; }
    uint32_t u = f2u(x);
    e += (int)((u >> 23) & 0xffu) - 127;
    float m = u2f((u & 0x007fffffu) | 0x3f800000u);
    if (m > SQRT2F) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = s2 * (3.33333333333333333333e-1f + s2 * (2.0e-1f + s2 * (1.42857142857142857143e-1f +
              s2 * (1.11111111111111111111e-1f + s2 * 9.09090909090909090909e-2f))));
    float lm = (s + s * p) * 2.0f;
    float fe = (float)e;
    return fe * LN2_HI + (fe * LN2_LO + lm);
}

#define PI_F 3.14159265358979323846f
#define PI_2F 1.57079632679489661923f
#define PI_4F 0.78539816339744830962f
#define TAN_PI_8F 0.41421356237309504880f

static float atan01(float t)   /* t in [0, 1] */
{
    float off = 0.0f;
    if (t > TAN_PI_8F) { t = (t - 1.0f) / (t + 1.0f); off = PI_4F; }
    float t2 = t * t;
    float p = t2 * (-3.33333333333333333333e-1f + t2 * (2.0e-1f + t2 * (-1.42857142857142857143e-1f +
              t2 * (1.11111111111111111111e-1f + t2 * (-9.09090909090909090909e-2f + t2 * (7.69230769230769230769e-2f +
              t2 * (-6.66666666666666666667e-2f + t2 * (5.88235294117647058824e-2f + t2 * (-5.26315789473684210526e-2f +
              t2 * 4.76190476190476190476e-2f)))))))));
    return off + (t + t * p);
}

float orc_det_atan2f(float y, float x)
{
    if (x != x || y != y) return x + y;
    float ax = fabsf(x), ay = fabsf(y);
    int xneg = (f2u(x) & 0x80000000u) != 0;
    int yneg = (f2u(y) & 0x80000000u) != 0;
    float a;
    float inf = u2f(0x7f800000u);
    if (ay == 0.0f) {
        a = xneg ? PI_F : 0.0f;
    } else if (ax == 0.0f) {
        a = PI_2F;
    } else if (ax == inf && ay == inf) {
        a = xneg ? 3.0f * PI_4F : PI_4F;
    } else if (ax == inf) {
        a = xneg ? PI_F : 0.0f;
    } else if (ay == inf) {
        a = PI_2F;
    } else {
        if (ay <= ax) a = atan01(ay / ax);
        else a = PI_2F - atan01(ax / ay);
        if (xneg) a = PI_F - a;
    }
    return yneg ? -a : a;
}

float orc_det_powi(float x, int32_t n)
{
    uint32_t un = n < 0 ? (uint32_t)0 - (uint32_t)n : (uint32_t)n;
    float r = 1.0f, b = x;
    while (un) {
        if (un & 1u) r = r * b;
        un >>= 1;
        if (un) b = b * b;
    }
    return n < 0 ? 1.0f / r : r;
}

/* x^y of the shading formulas (compute.wgsl:4437, 4505, 4603, 446-480): exp(y log x) for
 * x > 0; 0^y = 0 (y > 0), 1 (y = 0), +inf (y < 0); 1^y = 1; x < 0 -> NaN (pow is undefined
 * there in GLSL/WGSL). */
float orc_det_powf(float x, float y)
{
    if (x != x || y != y) return x + y;
    if (x < 0.0f) return u2f(0x7fc00000u);
    if (x == 1.0f) return 1.0f;
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : u2f(0x7f800000u));
    return orc_det_expf(y * orc_det_logf(x));
}

/* acos x = atan2(sqrt((1 - x)(1 + x)), x) on [-1, 1] (SunIntensity, compute.wgsl:486) */
float orc_det_acosf(float x)
{
    if (x != x) return x;
    if (x > 1.0f || x < -1.0f) return u2f(0x7fc00000u);
    return orc_det_atan2f(sqrtf((1.0f - x) * (1.0f + x)), x);
}

float orc_det_pow045(float x)
{
    if (x != x) return x;
    if (x <= 0.0f) return 0.0f;
    if (x == 1.0f) return 1.0f;
    return orc_det_expf(0.449999988079071045f * orc_det_logf(x));
}
