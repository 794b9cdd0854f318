// tsi_rules.h -- forward and reverse rules of every tape form (compute.wgsl:3642-3766 and
// 3786-3963), shared by the tape interpreter and by the straight-line per-tape code, so that
// both evaluate a record with the same binary32 operations.
#pragma once
#include "tsi_interval.h"
#include "tsi_tape.h"

namespace tsi {

// Transcendental and rarely used forms.  The interpreter calls them through the __noinline__
// wrappers below (their per-lane branches then stay out of its dispatch loop, which compiles
// to uniform scalar branches only); straight-line per-tape code inlines them.
TSI_DI Iv rare_forward_i(uint32_t form, Iv l, Iv r, float c) {
    switch (form) {
    case F_SIN: return SIN_I(l);
    case F_COS: return COS_I(l);
    case F_EXP: return EXP_I(l);
    case F_LOG: return LOG_I(l);
    case F_FLOOR: return FLOOR_I(l);
    case F_POW_RR: return POW_II(l, r);
    case F_POW_RL: return POW_II(l, I_FROM(c));
    case F_MOD_RR: return MOD_II(l, r);
    case F_MOD_RL: return MOD_II(l, I_FROM(c));
    case F_ATAN2_RR: return ATAN2_II(l, r);
    default: return ATAN2_II(l, I_FROM(c));
    }
}

struct Iv2 { Iv a, b; };

__device__ __noinline__ Iv rare_forward(uint32_t form, Iv l, Iv r, float c) { return rare_forward_i(form, l, r, c); }

// Reverse rules of the same forms (compute.wgsl:2806-2970): .a for the left operand, .b for
// the right one of two-register forms.
TSI_DI Iv2 rare_reverse_i(uint32_t form, Iv l, Iv r, float c) {
    const Iv ONE = I_FROM(1.0f);
    Iv2 x{I_FROM(0.0f), I_FROM(0.0f)};
    switch (form) {
    case F_SIN: x.a = COS_I(l); break;
    case F_COS: x.a = NEG_I(SIN_I(l)); break;
    case F_EXP: x.a = EXP_I(l); break;
    case F_LOG: x.a = DIV_II(ONE, l); break;
    case F_FLOOR: x.a = FLOOR_I_RR(l); break;
    case F_POW_RL: x.a = MUL_II(I_FROM(c), POW_II(l, SUB_IF(I_FROM(c), 1.0f))); break;
    case F_ATAN2_RL: x.a = DIV_II(NEG_I(l), ADD_II(SQR_I(I_FROM(c)), SQR_I(l))); break;
    case F_POW_RR:
        x.a = MUL_II(r, POW_II(l, SUB_IF(r, 1.0f)));
        x.b = MUL_II(POW_II(l, r), LOG_I(l));
        break;
    default: {   // F_ATAN2_RR
        const Iv den = ADD_II(SQR_I(r), SQR_I(l));
        x.a = DIV_II(NEG_I(l), den);
        x.b = DIV_II(r, den);
        break;
    }
    }
    return x;
}
__device__ __noinline__ Iv2 rare_reverse(uint32_t form, Iv l, Iv r, float c) { return rare_reverse_i(form, l, r, c); }

// MIN_II_RR / MAX_II_RR (compute.wgsl:2884-2903) as a choice: 1 left only, 2 right only, 3 both
TSI_DI uint32_t min_choice(Iv l, Iv r) { return l.hi < r.lo ? 1u : (r.hi < l.lo ? 2u : 3u); }
TSI_DI uint32_t max_choice(Iv l, Iv r) { return l.hi < r.lo ? 2u : (r.hi < l.lo ? 1u : 3u); }
TSI_DI bool is_minmax(uint32_t form) {
    return form == F_MIN_RR || form == F_MAX_RR || form == F_MIN_RL || form == F_MAX_RL;
}

TSI_DI bool is_rare(uint32_t form) {
    switch (form) {
    case F_SIN: case F_COS: case F_EXP: case F_LOG: case F_FLOOR: case F_POW_RR: case F_POW_RL:
    case F_MOD_RR: case F_MOD_RL: case F_ATAN2_RR: case F_ATAN2_RL:
        return true;
    default:
        return false;
    }
}

// Forward rule of a form, before fix (compute.wgsl:3642-3766).  l / r are the register
// operands (r unused by one-register forms; for lit/reg forms the register is r), c the literal.
// INLINE_RARE: the rare forms inline (straight-line code) or call (interpreter).
template <bool INLINE_RARE>
TSI_DI Iv fwd_form_t(uint32_t form, Iv l, Iv r, float c) {
    switch (form) {
    case F_ADD_RR: return ADD_II(l, r);
    case F_ADD_RL: return ADD_IF(l, c);
    case F_SUB_RR: return SUB_II(l, r);
    case F_SUB_RL: return SUB_IF(l, c);
    case F_SUB_LR: return SUB_II(I_FROM(c), r);
    case F_MUL_RR: return MUL_II(l, r);
    case F_MUL_RL: return MUL_IF(l, c);
    case F_SQR: return SQR_I(l);
    case F_SQRT: return SQRT_I(l);
    case F_MIN_RR: return MIN_II(l, r);
    case F_MIN_RL: return MIN_IF(l, c);
    case F_MAX_RR: return MAX_II(l, r);
    case F_MAX_RL: return MAX_IF(l, c);
    case F_NEG: return NEG_I(l);
    case F_ABS: return ABS_I(l);
    case F_RET: return l;
    case F_DIV_RR: return DIV_II(l, r);
    case F_DIV_RL: return DIV_IF(l, c);
    case F_DIV_LR: return DIV_II(I_FROM(c), r);
    case F_SIN: case F_COS: case F_EXP: case F_LOG: case F_FLOOR: case F_POW_RR: case F_POW_RL:
    case F_MOD_RR: case F_MOD_RL: case F_ATAN2_RR: case F_ATAN2_RL:
        return INLINE_RARE ? rare_forward_i(form, l, r, c) : rare_forward(form, l, r, c);
    case F_NOP_RR: return I_FROM(c);
    case F_ZERO: return I_FROM(0.0f);
    default: return I_FROM(__uint_as_float(0xff800000u));
    }
}

// SQRT_I_RR (compute.wgsl:2878-2882) given v = SQRT_I(x): 1 / (2 v), the upper bound capped at
// 1e6.  The compiled engine passes the forward value of the record itself where its fix is the
// identity (then the same bits as SQRT_I of the saved operand), so the reverse pass runs no
// square root.
TSI_DI Iv sqrt_rr_of_value(Iv v) {
    Iv d = DIV_II(I_FROM(1.0f), MUL_IF(v, 2.0f));
    d.hi = wmin(d.hi, kBig);
    return d;
}

// Reverse rule of a form, before fix (compute.wgsl:2806-2970 as called from 3786-3963): the
// derivative with respect to the first register operand in .a (`x` the saved value of it) and
// to the second in .b (`y`, two-register forms).  MIN/MAX give the undecided choice rules
// MIN_II_RR / MAX_II_RR (2884-2903).
template <bool INLINE_RARE>
TSI_DI Iv2 rev_form_t(uint32_t form, Iv x, Iv y, float c) {
    const Iv Z = I_FROM(0.0f), ONE = I_FROM(1.0f), NINF = I_FROM(__uint_as_float(0xff800000u));
    Iv2 d{Z, Z};
    switch (form) {
    // one register operand (side 1)
    case F_SQR: d.a = MUL_IF(x, 2.0f); break;
    case F_SQRT: d.a = sqrt_rr_of_value(SQRT_I(x)); break;
    case F_NEG: d.a = I_FROM(-1.0f); break;
    case F_ADD_RL: case F_SUB_RL: case F_RET: case F_MOD_RL: d.a = ONE; break;
    case F_MUL_RL: d.a = I_FROM(c); break;
    case F_DIV_RL: d.a = DIV_II(ONE, I_FROM(c)); break;
    case F_ABS: d.a = ABS_I_RR(x); break;
    case F_MIN_RL: d.a = sel(x.hi < c, ONE, sel(c < x.lo, Z, ONE)); break;
    case F_MAX_RL: d.a = sel(x.hi < c, Z, ONE); break;
    case F_SIN: case F_COS: case F_EXP: case F_LOG: case F_FLOOR: case F_POW_RL: case F_ATAN2_RL:
        d.a = (INLINE_RARE ? rare_reverse_i(form, x, Z, c) : rare_reverse(form, x, Z, c)).a;
        break;
    case F_NINF_RL: d.a = NINF; break;
    // two register operands (side 0)
    case F_ADD_RR: d.a = ONE; d.b = ONE; break;
    case F_SUB_RR: d.a = ONE; d.b = I_FROM(-1.0f); break;
    case F_MUL_RR: d.a = y; d.b = x; break;
    case F_MIN_RR: {
        const bool lt = x.hi < y.lo, gt = y.hi < x.lo;
        d.a = sel(lt, ONE, sel(gt, Z, ONE));
        d.b = sel(lt, Z, ONE);
        break;
    }
    case F_MAX_RR: {
        const bool lt = x.hi < y.lo, gt = y.hi < x.lo;
        d.a = sel(lt, Z, ONE);
        d.b = sel(lt, ONE, sel(gt, Z, ONE));
        break;
    }
    case F_DIV_RR: d.a = DIV_II(ONE, y); d.b = NEG_I(DIV_II(x, SQR_I(y))); break;
    case F_POW_RR: case F_ATAN2_RR:
        d = INLINE_RARE ? rare_reverse_i(form, x, y, c) : rare_reverse(form, x, y, c);
        break;
    case F_NOP_RR: break;
    case F_NINF_RR: d.a = NINF; d.b = NINF; break;
    // literal then register (side 2): x is the register operand
    case F_SUB_LR: d.a = I_FROM(-1.0f); break;
    case F_DIV_LR: d.a = NEG_I(DIV_II(I_FROM(c), SQR_I(x))); break;
    case F_NINF_LR: d.a = NINF; break;
    default: break;
    }
    return d;
}

TSI_DI Iv fwd_form(uint32_t form, Iv l, Iv r, float c) { return fwd_form_t<false>(form, l, r, c); }
TSI_DI Iv2 rev_form(uint32_t form, Iv x, Iv y, float c) { return rev_form_t<false>(form, x, y, c); }

}  // namespace tsi
