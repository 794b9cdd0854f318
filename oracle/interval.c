/*
 * interval.c -- forward interval evaluation, reverse AD with MIN/MAX choices and choice
 * reachability of the oracle.  TEST INFRASTRUCTURE ONLY (see tsi_oracle.h).
 *
 * Each function restates the compute.wgsl function of the same name (line cited).
 * Intervals are not outward rounded (the reference does not round outward either).
 */
#include "tsi_oracle.h"
#include <math.h>
#include <string.h>

#define BIG 1000000.0f                       /* compute.wgsl:2534 */
#define PI_W 3.141592653589793f              /* compute.wgsl:2477 */
#define TWOPI_W 6.283185307179586f           /* compute.wgsl:2478 */
#define HALFPI_W 1.5707963267948966f         /* compute.wgsl:2479 */

typedef orc_iv iv;

static inline float wmin(float a, float b) { return orc_wmin(a, b); }
static inline float wmax(float a, float b) { return orc_wmax(a, b); }
static inline iv I(float lo, float hi) { iv r; r.lo = lo; r.hi = hi; return r; }
static inline iv I_FROM(float a) { return I(a, a); }                                          /* 2486 */
static inline iv I_FROM_UNORDERED(float a, float b) { return I(wmin(a, b), wmax(a, b)); }    /* 2494 */
static inline int I_CONTAINS_ZERO(iv a) { return a.lo <= 0.0f && 0.0f <= a.hi; }             /* 2498 */
static inline int is_nan_w(float a) { return !(a < 0.0f || a > 0.0f || a == 0.0f); }         /* 2672 */
static inline float DIAMETER(iv a) { return a.hi - a.lo; }                                    /* 2506 */
static inline iv ADD_II(iv a, iv b) { return I(a.lo + b.lo, a.hi + b.hi); }                  /* 2539 */
static inline iv ADD_IF(iv a, float b) { return I(a.lo + b, a.hi + b); }                     /* 2543 */
static inline iv SUB_II(iv a, iv b) { return I(a.lo - b.hi, a.hi - b.lo); }                  /* 2547 */
static inline iv SUB_IF(iv a, float b) { return I(a.lo - b, a.hi - b); }                     /* 2551 */
static inline iv NEG_I(iv a) { return I(-a.hi, -a.lo); }                                     /* 2580 */

static iv MUL_II(iv a, iv b)                                                                  /* 2555 */
{
    float x = a.lo * b.lo, y = a.lo * b.hi, z = a.hi * b.lo, w = a.hi * b.hi;
    return I(wmin(wmin(x, y), wmin(z, w)), wmax(wmax(x, y), wmax(z, w)));
}

static iv MUL_IF(iv a, float b)                                                               /* 2563 */
{
    if (b >= 0.0f) return I(a.lo * b, a.hi * b);
    return I(a.hi * b, a.lo * b);
}

static iv SQRT_I(iv a) { return I(sqrtf(wmax(a.lo, 0.0f)), sqrtf(wmax(a.hi, 0.0f))); }          /* 2572 */
static iv EXP_I(iv a) { return I(orc_det_expf(a.lo), orc_det_expf(a.hi)); }                     /* 2576 */
static iv LOG_I(iv a) { return I(orc_det_logf(wmax(a.lo, 0.0f)), orc_det_logf(wmax(a.hi, 0.0f))); } /* 2584 */

static iv DIV_II(iv a, iv b)                                                                  /* 2598 */
{
    float x = a.lo / b.lo, y = a.lo / b.hi, z = a.hi / b.lo, w = a.hi / b.hi;
    return I(wmin(wmin(x, y), wmin(z, w)), wmax(wmax(x, y), wmax(z, w)));
}

static inline iv DIV_IF(iv a, float b) { return DIV_II(a, I_FROM(b)); }                       /* 2609 */
static inline iv MIN_II(iv a, iv b) { return I(wmin(a.lo, b.lo), wmin(a.hi, b.hi)); }        /* 2613 */
static inline iv MIN_IF(iv a, float b) { return I(wmin(a.lo, b), wmin(a.hi, b)); }           /* 2617 */
static inline iv MAX_IF(iv a, float b) { return I(wmax(a.lo, b), wmax(a.hi, b)); }           /* 2621 */
static inline iv MAX_II(iv a, iv b) { return I(wmax(a.lo, b.lo), wmax(a.hi, b.hi)); }        /* 2625 */

static iv ABS_I(iv a)                                                                         /* 2629 */
{
    if (I_CONTAINS_ZERO(a)) return I(0.0f, wmax(-a.lo, a.hi));
    if (a.hi < 0.0f) return I(-a.hi, -a.lo);
    return a;
}

static iv ABS_I_RR(iv a)                                                                      /* 2640 */
{
    if (I_CONTAINS_ZERO(a)) return I(-1.0f, 1.0f);
    if (a.hi < 0.0f) return I_FROM(-1.0f);
    return I_FROM(1.0f);
}

static iv I_EMPTY(void) { return I(INFINITY, -INFINITY); }                                   /* 2655 */

static iv CAP_II(iv a, iv b)                                                                  /* 2664 */
{
    if (a.hi < b.lo || b.hi < a.lo) return I_EMPTY();
    return I(wmax(a.lo, b.lo), wmin(a.hi, b.hi));
}

static iv SQR_I(iv a)                                                                         /* 2689 */
{
    float x = a.lo * a.lo, y = a.hi * a.hi;
    if (I_CONTAINS_ZERO(a)) return I(0.0f, wmax(x, y));
    return I(wmin(x, y), wmax(x, y));
}

static iv FLOOR_I(iv a) { return I(floorf(a.lo), floorf(a.hi)); }                            /* 2704 */

static iv FLOOR_I_RR(iv a)                                                                    /* 2708 */
{
    if (floorf(a.hi) != floorf(a.lo)) return I(-BIG, BIG);
    return I_FROM(0.0f);
}

/* WGSL atan2(y, x) */
static inline float watan2(float y, float x) { return orc_det_atan2f(y, x); }

static iv ATAN2_II(iv x, iv y)                                                                /* 2724 */
{
    iv range_atan = I(-PI_W, PI_W);
    iv half_range_atan = I(-HALFPI_W, HALFPI_W);
    if (y.lo == 0.0f) return I(0.0f, y.hi);
    if (y.hi == 0.0f) return I(y.lo, 0.0f);
    if (x.lo == 0.0f && x.hi == 0.0f) {
        if (y.lo == 0.0f && y.hi == 0.0f) return I_EMPTY();
        if (y.lo >= 0.0f) return I_FROM(HALFPI_W);
        if (y.hi <= 0.0f) return I_FROM(-HALFPI_W);
        return half_range_atan;
    } else if (x.lo > 0.0f) {
        if (y.lo == 0.0f && y.hi == 0.0f) return y;
        if (y.lo >= 0.0f) return I(watan2(y.lo, x.hi), watan2(y.hi, x.lo));
        if (y.hi <= 0.0f) return I(watan2(y.lo, x.lo), watan2(y.hi, x.hi));
        return I(watan2(y.lo, x.lo), watan2(y.hi, x.lo));
    } else if (x.hi < 0.0f) {
        if (y.lo == 0.0f && y.hi == 0.0f) return I_FROM(PI_W);
        if (y.lo >= 0.0f) return I(watan2(y.hi, x.hi), watan2(y.lo, x.lo));
        if (y.hi < 0.0f) return I(watan2(y.hi, x.lo), watan2(y.lo, x.hi));
        return range_atan;
    } else {
        if (x.lo == 0.0f) {
            if (y.lo == 0.0f && y.hi == 0.0f) return y;
            if (y.lo >= 0.0f) return I(watan2(y.lo, x.hi), half_range_atan.hi);
            if (y.hi <= 0.0f) return I(half_range_atan.lo, watan2(y.hi, x.hi));
            return half_range_atan;
        } else if (x.hi == 0.0f) {
            if (y.lo == 0.0f && y.hi == 0.0f) return I_FROM(PI_W);
            if (y.lo >= 0.0f) return I(HALFPI_W, watan2(y.lo, x.lo));
            if (y.hi < 0.0f) return I(watan2(y.hi, x.lo), -HALFPI_W);
            return range_atan;
        } else {
            if (y.lo >= 0.0f) return I(watan2(y.lo, x.hi), watan2(y.lo, x.lo));
            if (y.hi < 0.0f) return I(watan2(y.hi, x.lo), watan2(y.hi, x.hi));
            return range_atan;
        }
    }
}

static void ATAN2_II_RR(iv l, iv r, iv out[2])                                                /* 2805 */
{
    iv den = ADD_II(SQR_I(r), SQR_I(l));
    out[0] = DIV_II(NEG_I(l), den);
    out[1] = DIV_II(r, den);
}

static iv SQR_I_RR(iv l) { return MUL_IF(l, 2.0f); }                                         /* 2813 */

static iv MOD_II(iv a, iv b) { return SUB_II(a, MUL_II(b, FLOOR_I(DIV_II(a, b)))); }          /* 2818 */

static void MOD_II_RR(iv l, iv r, iv out[2])                                                  /* 2822 */
{
    iv a = FLOOR_I(DIV_II(l, r));
    out[0] = I_FROM(1.0f);
    out[1] = NEG_I(ADD_II(a, MUL_II(r, FLOOR_I_RR(a))));
}

/* WGSL i32(f32): NaN -> 0, saturating */
static int32_t f2i_sat(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}

static iv POW_II(iv a, iv b)                                                                  /* 2828 */
{
    int32_t power = f2i_sat(rintf(b.lo));
    float lower, higher;
    if (a.lo <= 0.0f) {
        if (power % 2 == 0) lower = orc_det_powi(-a.lo, power);
        else lower = -orc_det_powi(-a.lo, power);
    } else {
        lower = orc_det_powi(a.lo, power);
    }
    if (a.hi <= 0.0f) {
        if (power % 2 == 0) higher = orc_det_powi(-a.hi, power);
        else higher = -orc_det_powi(-a.hi, power);
    } else {
        higher = orc_det_powi(a.hi, power);
    }
    if (I_CONTAINS_ZERO(a)) return I(0.0f, wmax(lower, higher));
    return I(wmin(lower, higher), wmax(lower, higher));
}

static void POW_II_RR(iv l, iv r, iv out[2])                                                  /* 2861 */
{
    out[0] = MUL_II(r, POW_II(l, SUB_IF(r, 1.0f)));
    out[1] = MUL_II(POW_II(l, r), LOG_I(l));
}

static iv LOG_I_RR(iv l) { return DIV_II(I_FROM(1.0f), l); }                                 /* 2865 */
static iv EXP_I_RR(iv l) { return EXP_I(l); }                                                /* 2869 */

static iv SQRT_I_RR(iv l)                                                                     /* 2877 */
{
    iv res = DIV_II(I_FROM(1.0f), MUL_IF(SQRT_I(l), 2.0f));
    res.hi = wmin(res.hi, BIG);
    return res;
}

static void MIN_II_RR(iv l, iv r, iv out[2])                                                  /* 2884 */
{
    if (l.hi < r.lo) { out[0] = I_FROM(1.0f); out[1] = I_FROM(0.0f); }
    else if (r.hi < l.lo) { out[0] = I_FROM(0.0f); out[1] = I_FROM(1.0f); }
    else { out[0] = I_FROM(1.0f); out[1] = I_FROM(1.0f); }
}

static void MAX_II_RR(iv l, iv r, iv out[2])                                                  /* 2894 */
{
    if (l.hi < r.lo) { out[0] = I_FROM(0.0f); out[1] = I_FROM(1.0f); }
    else if (r.hi < l.lo) { out[0] = I_FROM(1.0f); out[1] = I_FROM(0.0f); }
    else { out[0] = I_FROM(1.0f); out[1] = I_FROM(1.0f); }
}

static void DIV_II_RR(iv l, iv r, iv out[2])                                                  /* 2908 */
{
    out[0] = DIV_II(I_FROM(1.0f), r);
    out[1] = NEG_I(DIV_II(l, SQR_I(r)));
}

static float mymod(float x, float y) { return x - y * floorf(x / y); }                        /* 2920 */

static iv SIN_I(iv a)                                                                         /* 2926 */
{
    iv whole = I(-1.0f, 1.0f);
    if (DIAMETER(a) > TWOPI_W) return whole;
    float lq = floorf(a.lo / HALFPI_W);
    float hq = floorf(a.hi / HALFPI_W);
    if (hq - lq > 4.0f) return whole;
    lq = mymod(lq, 4.0f);
    hq = mymod(hq, 4.0f);
    if (lq == hq) {
        if (a.hi - a.lo > PI_W) return whole;
        return I_FROM_UNORDERED(orc_det_sinf(a.lo), orc_det_sinf(a.hi));
    } else if ((lq == 3.0f && hq == 0.0f) || (lq == 1.0f && hq == 2.0f)) {
        return I_FROM_UNORDERED(orc_det_sinf(a.lo), orc_det_sinf(a.hi));
    } else if ((lq == 0.0f || lq == 3.0f) && (hq == 1.0f || hq == 2.0f)) {
        return I(wmin(orc_det_sinf(a.lo), orc_det_sinf(a.hi)), 1.0f);
    } else if ((lq == 1.0f || lq == 2.0f) && (hq == 3.0f || hq == 0.0f)) {
        return I(-1.0f, wmax(orc_det_sinf(a.lo), orc_det_sinf(a.hi)));
    }
    return whole;
}

static iv COS_I(iv a) { return SIN_I(ADD_II(a, I_FROM(HALFPI_W))); }                         /* 2959 */

static iv fix(iv x)                                                                           /* 3462 */
{
    float lo = x.lo, hi = x.hi;
    if (is_nan_w(x.hi)) hi = BIG;
    if (is_nan_w(x.lo)) lo = -BIG;
    return I(wmax(lo, -BIG), wmin(hi, BIG));
}

enum { OP_NOP = 0, OP_ADD = 5, OP_SUB = 6, OP_MUL = 7, OP_DIV = 8, OP_EXP = 9, OP_LOG = 10, OP_SQR = 11,
       OP_SQRT = 12, OP_POW = 13, OP_SIN = 14, OP_COS = 15, OP_NEG = 16, OP_MIN = 17, OP_MAX = 18,
       OP_ABS = 19, OP_FLOOR = 25, OP_RET = 28, OP_MOD = 29, OP_ATAN2 = 30 };

/* compute.wgsl:3642-3768 (one forward step, including fix) */
orc_iv orc_forward_op(uint32_t op, uint32_t side, orc_iv l, orc_iv r, float lit)
{
    iv val = I_FROM(0.0f);
    iv ninf = I_FROM(-INFINITY);
    switch (side) {
    case 1:
        switch (op) {
        case OP_SQR: val = SQR_I(l); break;
        case OP_SQRT: val = SQRT_I(l); break;
        case OP_EXP: val = EXP_I(l); break;
        case OP_LOG: val = LOG_I(l); break;
        case OP_NEG: val = NEG_I(l); break;
        case OP_ABS: val = ABS_I(l); break;
        case OP_FLOOR: val = FLOOR_I(l); break;
        case OP_SIN: val = SIN_I(l); break;
        case OP_COS: val = COS_I(l); break;
        case OP_ADD: val = ADD_IF(l, lit); break;
        case OP_SUB: val = SUB_IF(l, lit); break;
        case OP_DIV: val = DIV_IF(l, lit); break;
        case OP_MUL: val = MUL_IF(l, lit); break;
        case OP_MIN: val = MIN_IF(l, lit); break;
        case OP_MAX: val = MAX_IF(l, lit); break;
        case OP_POW: val = POW_II(l, I_FROM(lit)); break;
        case OP_RET: val = l; break;
        case OP_MOD: val = MOD_II(l, I_FROM(lit)); break;
        case OP_ATAN2: val = ATAN2_II(l, I_FROM(lit)); break;
        default: val = ninf; break;
        }
        break;
    case 0:
        switch (op) {
        case OP_ADD: val = ADD_II(l, r); break;
        case OP_SUB: val = SUB_II(l, r); break;
        case OP_DIV: val = DIV_II(l, r); break;
        case OP_MUL: val = MUL_II(l, r); break;
        case OP_MIN: val = MIN_II(l, r); break;
        case OP_MAX: val = MAX_II(l, r); break;
        case OP_POW: val = POW_II(l, r); break;
        case OP_MOD: val = MOD_II(l, r); break;
        case OP_ATAN2: val = ATAN2_II(l, r); break;
        case OP_NOP: val = I_FROM(lit); break;
        default: val = ninf; break;
        }
        break;
    case 2:
        switch (op) {
        case OP_SUB: val = SUB_II(I_FROM(lit), r); break;
        case OP_DIV: val = DIV_II(I_FROM(lit), r); break;
        default: val = ninf; break;
        }
        break;
    default:
        break;
    }
    return fix(val);
}

/* Algorithmic f32 op counts per executed instruction (forward incl. fix; reverse incl.
 * fix + MUL_II + ADD_II per receiving argument).  Compares, selects, sign flips and NaN
 * tests count 0. */
uint32_t orc_fwd_cost(uint32_t op, uint32_t side)
{
    uint32_t c = 0;
    switch (op) {
    case OP_ADD: case OP_SUB: c = 2; break;
    case OP_MUL: c = side == 0 ? 10 : 2; break;
    case OP_DIV: c = 10; break;
    case OP_SQR: c = 4; break;
    case OP_SQRT: c = 4; break;
    case OP_MIN: case OP_MAX: c = 2; break;
    case OP_ABS: c = 1; break;
    case OP_EXP: c = 2; break;
    case OP_LOG: c = 4; break;
    case OP_SIN: c = 16; break;
    case OP_COS: c = 17; break;
    case OP_POW: c = 8; break;
    case OP_MOD: c = 22; break;
    case OP_ATAN2: c = 8; break;
    default: c = 0; break;
    }
    if (side == 3) c = 0;
    return c + 2;
}

uint32_t orc_rev_cost(uint32_t op, uint32_t side)
{
    uint32_t rr = 0, nargs = side == 0 ? 2u : (side == 3 ? 0u : 1u);
    if (side == 3) return 0;
    switch (op) {
    case OP_SQR: rr = 2; break;
    case OP_SQRT: rr = 17; break;
    case OP_DIV: rr = 24; break;
    case OP_EXP: rr = 2; break;
    case OP_LOG: rr = 10; break;
    case OP_SIN: rr = 17; break;
    case OP_COS: rr = 16; break;
    case OP_POW: rr = 24; break;
    case OP_ATAN2: rr = 28; break;
    default: rr = 0; break;
    }
    return rr + 14u * nargs;
}

static inline uint32_t U_OP(const orc_instr *t) { return (t->opcode_uses >> 16) & ~(3u << 14); }
static inline uint32_t U_SIDE(const orc_instr *t) { return (t->opcode_uses >> 16) >> 14; }
static inline uint32_t U_USES(const orc_instr *t) { return t->opcode_uses & 0xFFFFu; }
static inline uint32_t U_L(const orc_instr *t) { return t->larg_rarg >> 16; }
static inline uint32_t U_R(const orc_instr *t) { return t->larg_rarg & 0xFFFFu; }

static inline int iv_zero(iv a) { return a.lo == 0.0f && a.hi == 0.0f; }

/* compute.wgsl:3622-3964 for one cell. */
orc_iv orc_eval_tape(const orc_instr *tape, uint32_t n, const orc_iv args[4], int take_derivative,
                     orc_iv *vals, orc_iv *sens, uint64_t *fwd_ops, uint64_t *rev_ops)
{
    iv *v = vals, *s = sens;
    for (uint32_t k = 0; k < 4 && k < n; k++) { v[k] = args[k]; s[k] = I_FROM(0.0f); }
    uint64_t fo = 0, ro = 0;
    for (uint32_t j = 4; j < n; j++) {                                  /* 3633 */
        const orc_instr *t = &tape[j];
        uint32_t op = U_OP(t), side = U_SIDE(t), li = U_L(t), ri = U_R(t);
        iv l = (side == 0 || side == 1) ? v[li] : I_FROM(0.0f);
        iv r = (side == 0 || side == 2) ? v[ri] : I_FROM(0.0f);
        v[j] = orc_forward_op(op, side, l, r, t->literal);
        s[j] = I_FROM(0.0f);
        fo += orc_fwd_cost(op, side);
    }
    iv field = v[n - 1];                                                /* 3774 */
    s[n - 1] = I_FROM(1.0f);                                            /* 3777 */
    if (take_derivative) {
        for (uint32_t j = n - 1; j >= 4; j--) {                         /* 3786 */
            const orc_instr *t = &tape[j];
            uint32_t op = U_OP(t), side = U_SIDE(t), li = U_L(t), ri = U_R(t);
            uint32_t luses = U_USES(&tape[li]), ruses = U_USES(&tape[ri]);
            iv sj = s[j];
            if (side == 1) {
                iv l = v[li];
                float lit = t->literal;
                if (iv_zero(sj) && luses == 1u) {
                    s[li] = I_FROM(0.0f);
                } else {
                    iv x, pr[2];
                    switch (op) {
                    case OP_SQR: x = SQR_I_RR(l); break;
                    case OP_SQRT: x = SQRT_I_RR(l); break;
                    case OP_NEG: x = I_FROM(-1.0f); break;
                    case OP_SIN: x = COS_I(l); break;
                    case OP_COS: x = NEG_I(SIN_I(l)); break;
                    case OP_ADD: x = I_FROM(1.0f); break;
                    case OP_SUB: x = I_FROM(1.0f); break;
                    case OP_MUL: x = I_FROM(lit); break;
                    case OP_DIV: DIV_II_RR(l, I_FROM(lit), pr); x = pr[0]; break;
                    case OP_EXP: x = EXP_I_RR(l); break;
                    case OP_LOG: x = LOG_I_RR(l); break;
                    case OP_FLOOR: x = FLOOR_I_RR(l); break;
                    case OP_ABS: x = ABS_I_RR(l); break;
                    case OP_MIN: MIN_II_RR(l, I_FROM(lit), pr); x = pr[0]; break;
                    case OP_MAX: MAX_II_RR(l, I_FROM(lit), pr); x = pr[0]; break;
                    case OP_POW: POW_II_RR(l, I_FROM(lit), pr); x = pr[0]; break;
                    case OP_RET: x = I_FROM(1.0f); break;
                    case OP_MOD: MOD_II_RR(l, I_FROM(lit), pr); x = pr[0]; break;
                    case OP_ATAN2: ATAN2_II_RR(l, I_FROM(lit), pr); x = pr[0]; break;
                    default: x = I_FROM(-INFINITY); break;
                    }
                    x = fix(x);
                    x = MUL_II(x, sj);
                    x = ADD_II(x, s[li]);
                    s[li] = x;
                    ro += orc_rev_cost(op, side);
                }
            } else if (side == 0) {
                iv l = v[li], r = v[ri];
                if (iv_zero(sj) && luses == 1u && ruses == 1u) {
                    s[li] = I_FROM(0.0f);
                    s[ri] = I_FROM(0.0f);
                } else {
                    iv pr[2];
                    switch (op) {
                    case OP_MIN: MIN_II_RR(l, r, pr); break;
                    case OP_MUL: pr[0] = r; pr[1] = l; break;
                    case OP_DIV: DIV_II_RR(l, r, pr); break;
                    case OP_ADD: pr[0] = I_FROM(1.0f); pr[1] = I_FROM(1.0f); break;
                    case OP_SUB: pr[0] = I_FROM(1.0f); pr[1] = I_FROM(-1.0f); break;
                    case OP_MAX: MAX_II_RR(l, r, pr); break;
                    case OP_POW: POW_II_RR(l, r, pr); break;
                    case OP_ATAN2: ATAN2_II_RR(l, r, pr); break;
                    case OP_NOP: pr[0] = I_FROM(0.0f); pr[1] = I_FROM(0.0f); break;
                    default: pr[0] = I_FROM(-INFINITY); pr[1] = I_FROM(-INFINITY); break;
                    }
                    pr[0] = fix(pr[0]);
                    pr[1] = fix(pr[1]);
                    pr[0] = MUL_II(pr[0], sj);
                    pr[1] = MUL_II(pr[1], sj);
                    pr[0] = ADD_II(pr[0], s[li]);
                    pr[1] = ADD_II(pr[1], s[ri]);
                    s[li] = pr[0];
                    s[ri] = pr[1];
                    ro += orc_rev_cost(op, side);
                }
            } else if (side == 2) {
                float lit = t->literal;
                iv r = v[ri];
                if (iv_zero(sj) && ruses == 1u) {
                    s[ri] = I_FROM(0.0f);
                } else {
                    iv x, pr[2];
                    switch (op) {
                    case OP_SUB: x = I_FROM(-1.0f); break;
                    case OP_DIV: DIV_II_RR(I_FROM(lit), r, pr); x = pr[1]; break;
                    default: x = I_FROM(-INFINITY); break;
                    }
                    x = fix(x);
                    x = MUL_II(x, sj);
                    x = ADD_II(x, s[ri]);
                    s[ri] = x;
                    ro += orc_rev_cost(op, side);
                }
            }
            if (j == 4) break;
        }
    }
    if (fwd_ops) *fwd_ops += fo;
    if (rev_ops) *rev_ops += ro;
    return field;
}

/* Row P: reachability from the output through choice edges (the MIN_II_RR / MAX_II_RR
 * rules of compute.wgsl:2884-2903 applied to this cell's forward values). */
void orc_prune(const orc_instr *tape, uint32_t n, const orc_iv *v, uint8_t *live, uint8_t *choice)
{
    memset(live, 0, n);
    memset(choice, 0, n);
    if (n == 0) return;
    live[n - 1] = 1;
    for (uint32_t j = n - 1; j >= 4; j--) {
        if (live[j]) {
            const orc_instr *t = &tape[j];
            uint32_t op = U_OP(t), side = U_SIDE(t), li = U_L(t), ri = U_R(t);
            uint32_t use_l = (side == 0 || side == 1), use_r = (side == 0 || side == 2);
            if ((op == OP_MIN || op == OP_MAX) && side != 3) {
                iv l = side == 2 ? I_FROM(t->literal) : v[li];
                iv r = side == 1 ? I_FROM(t->literal) : v[ri];
                iv pr[2];
                if (op == OP_MIN) MIN_II_RR(l, r, pr); else MAX_II_RR(l, r, pr);
                uint8_t c = (uint8_t)((pr[0].lo != 0.0f ? 1 : 0) | (pr[1].lo != 0.0f ? 2 : 0));
                choice[j] = c;
                if (use_l && (c & 1)) live[li] = 1;
                if (use_r && (c & 2)) live[ri] = 1;
            } else {
                if (use_l) live[li] = 1;
                if (use_r) live[ri] = 1;
            }
        }
        if (j == 4) break;
    }
    for (uint32_t k = 0; k < 4 && k < n; k++) live[k] = live[k] ? 1 : 0;
}
