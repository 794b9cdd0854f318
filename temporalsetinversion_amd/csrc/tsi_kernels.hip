// tsi_kernels.hip -- CDNA4 (gfx950) kernels of the temporal set-inversion path.
//
// The reference runs one persistent WGSL kernel fed by a lock-free broker queue, one
// parent voxel per 8-thread workgroup, with every lane keeping a private 131 KB tape in
// global memory (compute.wgsl:3513-4056).  Here one frame is a fixed sequence of launches
// over a level-synchronous frontier:
//   k_frame_begin            seed the root (buf_main / first-frame seeding)
//   k_eval<REV> x levels     one 64-lane wavefront = 8 parents x 8 child octants; the tape
//                            is interpreted wave-uniformly (instructions via scalar loads,
//                            live values in a uniformly indexed VGPR file), then the cell is
//                            classified; subdivision / dealloc outputs are compacted with
//                            wavefront ballots + one atomic per wave and list
//   k_dealloc x levels       the dealloc cascade of compute.wgsl:3533-3543 / 4036-4044
//   k_frame_end/commit       free-list merge
// Frontier lists are split by whether the children will need the reverse pass, so the
// forward-only kernel carries no reverse-mode code or stack traffic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tsi_interval.h"
#include "tsi_kernels.h"
#include "tsi_tape.h"

namespace tsi {

// ---------------------------------------------------------------- wave helpers

TSI_DI uint32_t lane_id() { return threadIdx.x & 63u; }
TSI_DI uint64_t ballot(bool p) { return __ballot(p); }
TSI_DI uint32_t popc64(uint64_t m) { return uint32_t(__popcll(m)); }
TSI_DI uint32_t rank_in(uint64_t m) {
    // number of set bits of m below this lane
    return uint32_t(__popcll(m & ((1ull << lane_id()) - 1ull)));
}
TSI_DI uint32_t bcast(uint32_t v, int src) { return uint32_t(__shfl(int(v), src, 64)); }

// Appends one element per lane where p holds to list[0 .. cap), count at *cnt.  One atomic per
// wavefront; returns the lane's position (0xFFFFFFFF if none, or on overflow, which is flagged).
TSI_DI uint32_t wave_append(bool p, uint32_t value, uint32_t *list, uint32_t *cnt, uint32_t cap, uint32_t *err,
                            uint32_t errbit) {
    uint64_t m = ballot(p);
    if (m == 0) return 0xFFFFFFFFu;
    uint32_t n = popc64(m), base = 0;
    if (lane_id() == 0) base = atomicAdd(cnt, n);
    base = bcast(base, 0);
    uint32_t at = 0xFFFFFFFFu;
    if (p) {
        at = base + rank_in(m);
        if (at < cap) list[at] = value;
        else { atomicOr(err, errbit); at = 0xFFFFFFFFu; }
    }
    return at;
}

TSI_DI uint64_t wave_sum64(uint64_t v) {
    for (int off = 32; off > 0; off >>= 1) {
        uint32_t lo = uint32_t(v), hi = uint32_t(v >> 32);
        lo = uint32_t(__shfl_xor(int(lo), off, 64));
        hi = uint32_t(__shfl_xor(int(hi), off, 64));
        v += (uint64_t(hi) << 32) | lo;
    }
    return v;
}

// ---------------------------------------------------------------- node state

TSI_DI uint32_t meta_of(float4 s) { return __float_as_uint(s.w); }
constexpr uint32_t kFastBit = 0x80000000u;
constexpr uint32_t kChildMask = 0x7FFFFFFFu;   // child block + 1, 0 = no children

// ---------------------------------------------------------------- interpreter

// The interpreter's register file: kSlots intervals per lane, indexed only with
// wave-uniform slot numbers so the compiler keeps it in VGPRs (s_set_gpr_idx moves).
typedef float Slots[kSlots];

// Programs are read through the constant address space: the kernel never writes them, so the
// loads become scalar (s_load) and the slot numbers land in SGPRs (no per-lane waterfall).
typedef const __attribute__((address_space(4))) uint32_t *CWords;
// Transcendental and rarely used forms live behind a call: their per-lane branches stay out of
// the interpreter loop, whose dispatch then compiles to uniform scalar branches only.
__device__ __noinline__ Iv rare_forward(uint32_t form, Iv l, Iv r, float c) {
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

// Reverse rules of the same forms (compute.wgsl:2806-2970); returns the rule for the left
// operand in x0 and, for two-register forms, the right one in x1.
__device__ __noinline__ void rare_reverse(uint32_t form, Iv l, Iv r, float c, Iv *x0, Iv *x1) {
    const Iv ONE = I_FROM(1.0f);
    switch (form) {
    case F_SIN: *x0 = COS_I(l); break;
    case F_COS: *x0 = NEG_I(SIN_I(l)); break;
    case F_EXP: *x0 = EXP_I(l); break;
    case F_LOG: *x0 = DIV_II(ONE, l); break;
    case F_FLOOR: *x0 = FLOOR_I_RR(l); break;
    case F_POW_RL: *x0 = MUL_II(I_FROM(c), POW_II(l, SUB_IF(I_FROM(c), 1.0f))); break;
    case F_ATAN2_RL: *x0 = DIV_II(NEG_I(l), ADD_II(SQR_I(I_FROM(c)), SQR_I(l))); break;
    case F_POW_RR:
        *x0 = MUL_II(r, POW_II(l, SUB_IF(r, 1.0f)));
        *x1 = MUL_II(POW_II(l, r), LOG_I(l));
        break;
    default: {   // F_ATAN2_RR
        const Iv den = ADD_II(SQR_I(r), SQR_I(l));
        *x0 = DIV_II(NEG_I(l), den);
        *x1 = DIV_II(r, den);
        break;
    }
    }
}

// ---- pruned programs -------------------------------------------------------------------
// A program record's word 3 carries, in bits 16-31, the MIN/MAX choice each of the wave's 8
// parents decided (2 bits per parent: 1 = left operand only, 2 = right operand only,
// 3 = undecided).  Full programs carry 0xFFFF.  A parent's decision holds in its children
// (child box and t-window inside the parent's), so a decided MIN/MAX is a copy of its operand.
TSI_DI uint32_t lane_mode(uint32_t w3, uint32_t ps) { return (w3 >> (16u + 2u * ps)) & 3u; }
// MIN_II_RR / MAX_II_RR (compute.wgsl:2884-2903) as a choice: 1 left only, 2 right only, 3 both
TSI_DI uint32_t min_choice(Iv l, Iv r) { return l.hi < r.lo ? 1u : (r.hi < l.lo ? 2u : 3u); }
TSI_DI uint32_t max_choice(Iv l, Iv r) { return l.hi < r.lo ? 2u : (r.hi < l.lo ? 1u : 3u); }
TSI_DI bool is_minmax(uint32_t form) {
    return form == F_MIN_RR || form == F_MAX_RR || form == F_MIN_RL || form == F_MAX_RL;
}

// Per-lane record of the MIN/MAX choices made by a forward pass (2 bits per MIN/MAX record,
// 16 per word), kept in a per-wave buffer column for the liveness walk.
struct Choices {
    uint32_t *buf;   // this lane's column (stride 64 words)
    uint32_t word;
    uint32_t n;      // MIN/MAX records seen (wave-uniform)
};

// One forward step (compute.wgsl:3642-3772): w0 = {form, a, b, d}, w1 = {push, npush | side<<2,
// cost}, w2 = literal, w3 = {index, modes}.
template <bool PUSH, bool CHOICE>
TSI_DI void fwd_step(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t ps, Slots &Rlo, Slots &Rhi,
                     float2 *__restrict__ stk, Choices &ch) {
    const uint32_t form = w0 & 0xFFu, ia = (w0 >> 8) & 0xFFu, ib = (w0 >> 16) & 0xFFu, id = w0 >> 24;
    const Iv l = I(Rlo[ia], Rhi[ia]);
    const Iv r = I(Rlo[ib], Rhi[ib]);
    const float c = __uint_as_float(w2);
    Iv v;
    uint32_t choice = 3u;
    switch (form) {
    case F_ADD_RR: v = ADD_II(l, r); break;
    case F_ADD_RL: v = ADD_IF(l, c); break;
    case F_SUB_RR: v = SUB_II(l, r); break;
    case F_SUB_RL: v = SUB_IF(l, c); break;
    case F_SUB_LR: v = SUB_II(I_FROM(c), r); break;
    case F_MUL_RR: v = MUL_II(l, r); break;
    case F_MUL_RL: v = MUL_IF(l, c); break;
    case F_SQR: v = SQR_I(l); break;
    case F_SQRT: v = SQRT_I(l); break;
    case F_MIN_RR: {
        const uint32_t m = lane_mode(w3, ps);
        v = sel(m == 1u, l, sel(m == 2u, r, MIN_II(l, r)));
        choice = m != 3u ? m : min_choice(l, r);
        break;
    }
    case F_MIN_RL: {
        const uint32_t m = lane_mode(w3, ps);
        v = sel(m == 1u, l, sel(m == 2u, I_FROM(c), MIN_IF(l, c)));
        choice = m != 3u ? m : min_choice(l, I_FROM(c));
        break;
    }
    case F_MAX_RR: {
        const uint32_t m = lane_mode(w3, ps);
        v = sel(m == 1u, l, sel(m == 2u, r, MAX_II(l, r)));
        choice = m != 3u ? m : max_choice(l, r);
        break;
    }
    case F_MAX_RL: {
        const uint32_t m = lane_mode(w3, ps);
        v = sel(m == 1u, l, sel(m == 2u, I_FROM(c), MAX_IF(l, c)));
        choice = m != 3u ? m : max_choice(l, I_FROM(c));
        break;
    }
    case F_NEG: v = NEG_I(l); break;
    case F_ABS: v = ABS_I(l); break;
    case F_RET: v = l; break;
    case F_DIV_RR: v = DIV_II(l, r); break;
    case F_DIV_RL: v = DIV_IF(l, c); break;
    case F_DIV_LR: v = DIV_II(I_FROM(c), r); break;
    case F_SIN: case F_COS: case F_EXP: case F_LOG: case F_FLOOR: case F_POW_RR: case F_POW_RL:
    case F_MOD_RR: case F_MOD_RL: case F_ATAN2_RR: case F_ATAN2_RL:
        v = rare_forward(form, l, r, c);
        break;
    case F_NOP_RR: v = I_FROM(c); break;
    case F_ZERO: v = I_FROM(0.0f); break;
    default: v = I_FROM(__uint_as_float(0xff800000u)); break;
    }
    v = fix(v);
    Rlo[id] = v.lo;
    Rhi[id] = v.hi;
    if (CHOICE && is_minmax(form)) {
        ch.word |= choice << (2u * (ch.n & 15u));
        if ((ch.n & 15u) == 15u) {
            ch.buf[(ch.n >> 4) * 64u] = ch.word;
            ch.word = 0u;
        }
        ++ch.n;
    }
    if (PUSH) {
        const uint32_t npush = (w1 >> 16) & 3u;
        if (npush) {
            float2 *p = stk + size_t(w1 & 0xFFFFu) * 64;
            if (npush == 3) {
                p[0] = make_float2(r.lo, r.hi);
            } else {
                p[0] = make_float2(l.lo, l.hi);
                if (npush == 2) p[64] = make_float2(r.lo, r.hi);
            }
        }
    }
}

// Programs are fetched in 64-byte batches of four 16-byte records, one batch ahead of the
// batch being executed, so the scalar-load latency overlaps four interpreted instructions.
// Programs are padded with zero records to a multiple of four plus two batches.
constexpr uint32_t kBatch = 4;

// Forward pass (compute.wgsl:3633-3772); returns the algorithmic cost of the records run.
// PUSH saves the operand values the reverse rules need; CHOICE records MIN/MAX choices.
template <bool PUSH, bool CHOICE>
TSI_DI uint32_t run_forward(const FwdIns *prog, uint32_t nfwd, uint32_t ps, Slots &Rlo, Slots &Rhi,
                            float2 *__restrict__ stk, Choices &ch) {
    CWords base = (CWords)prog;
    uint32_t cost = 0;
    uint32_t cur[4 * kBatch];
#pragma unroll
    for (int i = 0; i < int(4 * kBatch); ++i) cur[i] = base[i];
    for (uint32_t k = 0; k < nfwd; k += kBatch) {
        uint32_t nxt[4 * kBatch];
        CWords nb = base + size_t(k + kBatch) * 4u;
#pragma unroll
        for (int i = 0; i < int(4 * kBatch); ++i) nxt[i] = nb[i];
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i)
            if (k + i < nfwd) {
                fwd_step<PUSH, CHOICE>(cur[4 * i], cur[4 * i + 1], cur[4 * i + 2], cur[4 * i + 3], ps, Rlo, Rhi, stk, ch);
                cost += cur[4 * i + 1] >> 24;
            }
#pragma unroll
        for (int i = 0; i < int(4 * kBatch); ++i) cur[i] = nxt[i];
    }
    if (CHOICE && (ch.n & 15u)) ch.buf[(ch.n >> 4) * 64u] = ch.word;
    return cost;
}

// One reverse step (compute.wgsl:3786-3963): w0 = {form, sj, sa, sb}, w1 = {flags, side, pop},
// w2 = literal, w3 = {cost, npush, modes}; ops_in = the saved operand values (left.lo, left.hi,
// right.lo, right.hi; a single saved interval sits in .xy).  `init` has a bit per sensitivity
// slot that holds an accumulated value; a slot read before its first write reads [+0, +0]
// (the I_FROM(0.0) every sens starts from), whichever instruction writes it first.
TSI_DI uint32_t rev_step(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t ps, Slots &Slo, Slots &Shi,
                         float4 ops_in, uint32_t &init) {
    const Iv Z = I_FROM(0.0f), ONE = I_FROM(1.0f), NINF = I_FROM(__uint_as_float(0xff800000u));
    const uint32_t form = w0 & 0xFFu, isj = (w0 >> 8) & 0xFFu, isa = (w0 >> 16) & 0xFFu, isb = w0 >> 24;
    const uint32_t flags = w1 & 0xFFu, side = (w1 >> 8) & 0xFFu;
    const float c = __uint_as_float(w2);
    const uint32_t cost = w3 & 0xFFu;
    Iv sj = Z;
    if (flags & R_SJ_ONE) {
        sj = ONE;
    } else if (!(flags & R_SJ_ZERO)) {
        if ((init >> isj) & 1u) sj = I(Slo[isj], Shi[isj]);
        init &= ~(1u << isj);
    }
    const bool sj0 = sj.lo == 0.0f && sj.hi == 0.0f;
    const bool fta = ((init >> isa) & 1u) == 0u, ftb = ((init >> isb) & 1u) == 0u;
    uint32_t ops = 0;
    if (side == 1) {
        const Iv l = I(ops_in.x, ops_in.y);
        uint32_t take = 1u;   // the left operand receives a contribution
        Iv x;
        switch (form) {
        case F_SQR: x = MUL_IF(l, 2.0f); break;
        case F_SQRT: {
            x = DIV_II(ONE, MUL_IF(SQRT_I(l), 2.0f));
            x.hi = wmin(x.hi, kBig);
            break;
        }
        case F_NEG: x = I_FROM(-1.0f); break;
        case F_SIN: case F_COS: case F_EXP: case F_LOG: case F_FLOOR: case F_POW_RL: case F_ATAN2_RL: {
            Iv unused;
            rare_reverse(form, l, Z, c, &x, &unused);
            break;
        }
        case F_ADD_RL: case F_SUB_RL: case F_RET: case F_MOD_RL: x = ONE; break;
        case F_MUL_RL: x = I_FROM(c); break;
        case F_DIV_RL: x = DIV_II(ONE, I_FROM(c)); break;
        case F_ABS: x = ABS_I_RR(l); break;
        case F_MIN_RL: {
            const uint32_t m = lane_mode(w3, ps);
            x = sel(m == 1u, ONE, sel(l.hi < c, ONE, sel(c < l.lo, Z, ONE)));
            take = m & 1u;
            break;
        }
        case F_MAX_RL: {
            const uint32_t m = lane_mode(w3, ps);
            x = sel(m == 1u, ONE, sel(l.hi < c, Z, ONE));
            take = m & 1u;
            break;
        }
        default: x = NINF; break;
        }
        const Iv old = fta ? Z : I(Slo[isa], Shi[isa]);
        Iv nv = take ? ADD_II(MUL_II(fix(x), sj), old) : old;
        const bool sc = sj0 && (flags & R_U1_A);
        nv = sel(sc, Z, nv);
        ops = sc ? 0u : cost;
        Slo[isa] = nv.lo;
        Shi[isa] = nv.hi;
        init |= 1u << isa;
    } else if (side == 0) {
        const Iv l = I(ops_in.x, ops_in.y), r = I(ops_in.z, ops_in.w);
        uint32_t takea = 1u, takeb = 1u;
        Iv x0, x1;
        switch (form) {
        case F_ADD_RR: x0 = ONE; x1 = ONE; break;
        case F_SUB_RR: x0 = ONE; x1 = I_FROM(-1.0f); break;
        case F_MUL_RR: x0 = r; x1 = l; break;
        case F_MIN_RR: {   // MIN_II_RR (compute.wgsl:2884), or the parent's decided choice
            const uint32_t m = lane_mode(w3, ps);
            const bool lt = l.hi < r.lo, gt = r.hi < l.lo;
            x0 = sel(m != 3u, ONE, sel(lt, ONE, sel(gt, Z, ONE)));
            x1 = sel(m != 3u, ONE, sel(lt, Z, ONE));
            takea = m & 1u;
            takeb = m >> 1;
            break;
        }
        case F_MAX_RR: {   // MAX_II_RR (compute.wgsl:2894), or the parent's decided choice
            const uint32_t m = lane_mode(w3, ps);
            const bool lt = l.hi < r.lo, gt = r.hi < l.lo;
            x0 = sel(m != 3u, ONE, sel(lt, Z, ONE));
            x1 = sel(m != 3u, ONE, sel(lt, ONE, sel(gt, Z, ONE)));
            takea = m & 1u;
            takeb = m >> 1;
            break;
        }
        case F_DIV_RR: x0 = DIV_II(ONE, r); x1 = NEG_I(DIV_II(l, SQR_I(r))); break;
        case F_POW_RR: case F_ATAN2_RR: rare_reverse(form, l, r, c, &x0, &x1); break;
        case F_NOP_RR: x0 = Z; x1 = Z; break;
        default: x0 = NINF; x1 = NINF; break;
        }
        const Iv olda = fta ? Z : I(Slo[isa], Shi[isa]);
        const Iv oldb = ftb ? Z : I(Slo[isb], Shi[isb]);
        Iv na = takea ? ADD_II(MUL_II(fix(x0), sj), olda) : olda;
        Iv nb = takeb ? ADD_II(MUL_II(fix(x1), sj), oldb) : oldb;
        const bool sc = sj0 && (flags & R_U1_A) && (flags & R_U1_B);
        na = sel(sc, Z, na);
        nb = sel(sc, Z, nb);
        ops = sc ? 0u : cost;
        Slo[isa] = na.lo;
        Shi[isa] = na.hi;
        Slo[isb] = nb.lo;
        Shi[isb] = nb.hi;
        init |= (1u << isa) | (1u << isb);
    } else if (side == 2) {
        const Iv r = I(ops_in.x, ops_in.y);
        Iv x;
        switch (form) {
        case F_SUB_LR: x = I_FROM(-1.0f); break;
        case F_DIV_LR: x = NEG_I(DIV_II(I_FROM(c), SQR_I(r))); break;
        default: x = NINF; break;
        }
        const Iv old = ftb ? Z : I(Slo[isb], Shi[isb]);
        Iv nv = ADD_II(MUL_II(fix(x), sj), old);
        const bool sc = sj0 && (flags & R_U1_B);
        nv = sel(sc, Z, nv);
        ops = sc ? 0u : cost;
        Slo[isb] = nv.lo;
        Shi[isb] = nv.hi;
        init |= 1u << isb;
    }
    return ops;
}

// Saved operands of one reverse record: npush (w3 bits 8-9) intervals at position pop.
TSI_DI float4 load_saved(uint32_t w1, uint32_t w3, const float2 *__restrict__ stk) {
    const uint32_t npush = (w3 >> 8) & 3u;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (npush) {
        const float2 *p = stk + size_t(w1 >> 16) * 64;
        const float2 q = p[0];
        v.x = q.x;
        v.y = q.y;
        if (npush == 2) {
            const float2 w = p[64];
            v.z = w.x;
            v.w = w.y;
        }
    }
    return v;
}

// Reverse pass (compute.wgsl:3780-3964) with sensitivities in the register file; returns the
// algorithmic op count of the non-short-circuited steps of this lane.  Two-deep pipeline: the
// saved operands of batch b+1 are loaded from the stack while batch b executes, and the
// program words of batch b+2 are fetched meanwhile.
TSI_DI uint32_t run_reverse(const RevIns *prog, uint32_t nrev, uint32_t ps, Slots &Slo, Slots &Shi,
                            const float2 *__restrict__ stk) {
    uint32_t ops = 0, init = 0xFu;   // the argument sens slots 0..3 start at [+0, +0]
#pragma unroll
    for (int k = 0; k < 4; ++k) { Slo[k] = 0.0f; Shi[k] = 0.0f; }
    CWords base = (CWords)prog;
    uint32_t cur[4 * kBatch], nxt[4 * kBatch];
    float4 scur[kBatch];
#pragma unroll
    for (int i = 0; i < int(4 * kBatch); ++i) cur[i] = base[i];
#pragma unroll
    for (int i = 0; i < int(4 * kBatch); ++i) nxt[i] = base[4 * kBatch + i];
#pragma unroll
    for (uint32_t i = 0; i < kBatch; ++i) scur[i] = load_saved(cur[4 * i + 1], cur[4 * i + 3], stk);
    for (uint32_t k = 0; k < nrev; k += kBatch) {
        float4 snxt[kBatch];
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) snxt[i] = load_saved(nxt[4 * i + 1], nxt[4 * i + 3], stk);
        uint32_t nn[4 * kBatch];
        CWords nb = base + size_t(k + 2 * kBatch) * 4u;
#pragma unroll
        for (int i = 0; i < int(4 * kBatch); ++i) nn[i] = nb[i];
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i)
            if (k + i < nrev)
                ops += rev_step(cur[4 * i], cur[4 * i + 1], cur[4 * i + 2], cur[4 * i + 3], ps, Slo, Shi, scur[i], init);
#pragma unroll
        for (int i = 0; i < int(4 * kBatch); ++i) { cur[i] = nxt[i]; nxt[i] = nn[i]; }
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) scur[i] = snxt[i];
    }
    return ops;
}

// Liveness of one cell: walks the program backwards from the output, following every
// register operand, and for MIN/MAX only the operands the cell's own choice keeps
// (compute.wgsl:2884-2903).  A bit per register-file slot says whether the value currently
// held there is still needed.  Writes the cell's liveness record (word w at rec[w * cap]):
// three bitmasks over the full program (live, left-kept, right-kept), 32 records per word.
TSI_DI void live_walk(const FwdIns *prog, uint32_t n, uint32_t out_slot, const Choices &ch, uint32_t nchunks,
                      uint32_t *rec, uint32_t cap) {
    CWords base = (CWords)prog;
    uint32_t m = 1u << out_slot;
    uint32_t lw = 0, Lw = 0, Rw = 0;
    int chunk = int(nchunks) - 1;
    uint32_t nmm = ch.n;
    uint32_t cw = 0;
    auto flush = [&](int c) {
        if (rec) {
            rec[size_t(kRecHead + c) * cap] = lw;
            rec[size_t(kRecHead + nchunks + c) * cap] = Lw;
            rec[size_t(kRecHead + 2 * nchunks + c) * cap] = Rw;
        }
        lw = Lw = Rw = 0;
    };
    const uint32_t rounded = (n + kBatch - 1) / kBatch * kBatch;
    for (int kb = int(rounded) - int(kBatch); kb >= 0; kb -= int(kBatch)) {
        uint32_t cur[4 * kBatch];
        CWords b = base + size_t(kb) * 4u;
#pragma unroll
        for (int i = 0; i < int(4 * kBatch); ++i) cur[i] = b[i];
#pragma unroll
        for (int i = int(kBatch) - 1; i >= 0; --i) {
            if (uint32_t(kb + i) >= n) continue;
            const uint32_t w0 = cur[4 * i], w1 = cur[4 * i + 1], w3 = cur[4 * i + 3];
            const uint32_t form = w0 & 0xFFu, ia = (w0 >> 8) & 0xFFu, ib = (w0 >> 16) & 0xFFu, id = w0 >> 24;
            const uint32_t side = (w1 >> 18) & 3u, idx = w3 & 0xFFFFu;
            while (chunk > int(idx >> 5)) flush(chunk--);
            const uint32_t live = (m >> id) & 1u;
            m &= ~(1u << id);
            uint32_t keepl = live, keepr = live;
            if (is_minmax(form)) {
                --nmm;
                if ((nmm & 15u) == 15u || nmm + 1u == ch.n) cw = ch.buf[(nmm >> 4) * 64u];
                const uint32_t c2 = (cw >> (2u * (nmm & 15u))) & 3u;
                keepl = live & c2;
                keepr = live & (c2 >> 1);
            }
            if ((side == 0 || side == 1) && keepl) m |= 1u << ia;
            if ((side == 0 || side == 2) && keepr) m |= 1u << ib;
            const uint32_t bit = 1u << (idx & 31u);
            lw |= live ? bit : 0u;
            Lw |= keepl ? bit : 0u;
            Rw |= keepr ? bit : 0u;
        }
    }
    while (chunk >= 0) flush(chunk--);
}

// ---------------------------------------------------------------- frame kernels

__global__ void k_frame_begin(KArgs a) {
    Counters *c = a.cnt;
    const int t = threadIdx.x;
    for (int i = t; i < kMaxLevels + 2; i += blockDim.x) {
        c->list_count[i][0] = 0;
        c->list_count[i][1] = 0;
        c->dlist_count[i] = 0;
    }
    for (int i = t; i < kNumStats; i += blockDim.x) c->stats[i] = 0;
    for (int i = t; i < kMaxLevels + 2; i += blockDim.x) c->pool_used[i] = 0;
    if (t == 0) {
        c->alloc_tickets = 0;
        c->pending = 0;
        c->frame_error = 0;
        if (c->root_pending) {
            c->root_pending = 0;
            const float4 s = a.st[0];
            const uint32_t meta = meta_of(s);
            const float4 lo = a.blo[0], hi = a.bhi[0];
            const float dx = (hi.x - lo.x) / 2.0f, dy = (hi.y - lo.y) / 2.0f, dz = (hi.z - lo.z) / 2.0f;
            const float w = wmax(dx, wmax(dy, dz));
            const bool fwd_only = !a.swept && ((meta & kChildMask) == 0 || (meta & kFastBit) || w <= a.epsilon);
            const int which = fwd_only ? 0 : 1;
            a.lists[0][which][0] = 0u;
            c->list_count[0][which] = 1;
        }
    }
}

template <bool REV>
__global__ void __launch_bounds__(256) k_eval(KArgs a, int level, float time) {
    const int which = REV ? 1 : 0;
    Counters *cnt = a.cnt;
    // a list that overflowed (flagged by wave_append) holds list_cap valid entries
    const uint32_t count = min(cnt->list_count[level][which], a.list_cap);
    if (count == 0) return;
    const uint32_t *list = a.lists[level & 1][which];
    uint32_t *next_f = a.lists[(level + 1) & 1][0];
    uint32_t *next_r = a.lists[(level + 1) & 1][1];
    const uint32_t lane = lane_id();
    const uint32_t ps = lane >> 3, o = lane & 7u;
    const uint32_t lx = o & 1u, ly = (o >> 1) & 1u, lz = o >> 2;
    // wave index made visibly uniform: the grid-stride loop and everything inside it then
    // compile as scalar control flow
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    float2 *stk = REV ? a.stack + size_t(wave) * a.stack_len * 64 + lane : nullptr;
    const float eps = a.epsilon;
    const bool prune = a.prune != 0;
    // this level's parents' liveness records (written by the previous level) and the next
    // level's (written here)
    const uint32_t *prec_in = a.prune_rec[level & 1][which];
    uint32_t *prec_f = a.prune_rec[(level + 1) & 1][0];
    uint32_t *prec_r = a.prune_rec[(level + 1) & 1][1];
    const __attribute__((address_space(4))) uint32_t *gdesc =
        (const __attribute__((address_space(4))) uint32_t *)a.gdesc[which];
    Choices ch;
    ch.buf = a.choice + size_t(wave) * a.choice_len * 64 + lane;

    uint64_t n_cells = 0, n_rev = 0, n_par = 0, n_enq = 0, rev_ops = 0, fwd_ops = 0, n_pruned = 0;
    Slots Rlo, Rhi;

    for (uint32_t g = wave; g * 8u < count; g += nwaves) {
        const uint32_t pidx = g * 8u + ps;
        const bool active = pidx < count;
        const uint32_t P = active ? list[pidx] : 0u;
        const float4 pst = a.st[P];
        const uint32_t pmeta = meta_of(pst);
        const bool isnew = active && (pmeta & kChildMask) == 0;
        const bool pfast = (pmeta & kFastBit) != 0;

        // ---- allocate child blocks for new parents: one atomic per wavefront ----
        uint32_t blk = (pmeta & kChildMask) - 1u;
        {
            const uint64_t req = ballot(isnew && o == 0);
            if (req) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&cnt->alloc_tickets, popc64(req));
                base = bcast(base, 0);
                uint32_t tk = base + rank_in(req);
                tk = bcast(tk, int(lane & ~7u));
                if (isnew) {
                    const uint32_t F = cnt->free_top;
                    if (tk < F) blk = a.free_stack[F - 1u - tk];
                    else blk = 1u + cnt->bump + (tk - F);
                }
            }
        }
        bool ok = active;
        if (isnew && blk >= a.nblocks) {
            ok = false;
            if (o == 0) atomicOr(&cnt->frame_error, kErrPool);
        }
        if (isnew && ok && o == 0) a.st[P].w = __uint_as_float((pmeta & kFastBit) | (blk + 1u));
        n_par += (active && o == 0) ? 1 : 0;

        const uint32_t C = ok ? blk * 8u + o : 0u;
        float4 cst = make_float4(1.0f, 0.0f, 0.0f, 0.0f);   // mallocV (compute.wgsl:3437-3449)
        if (ok && !isnew) cst = a.st[C];
        uint32_t cmeta = meta_of(cst);

        // ---- child box, t-window, point sampling (compute.wgsl:3582-3626) ----
        const float4 plo = a.blo[P], phi = a.bhi[P];
        const float dx = (phi.x - plo.x) / 2.0f, dy = (phi.y - plo.y) / 2.0f, dz = (phi.z - plo.z) / 2.0f;
        const float width = wmax(dx, wmax(dy, dz));
        Iv cx = I(plo.x + dx * float(lx), plo.x + dx * float(1u + lx));
        Iv cy = I(plo.y + dy * float(ly), plo.y + dy * float(1u + ly));
        Iv cz = I(plo.z + dz * float(lz), plo.z + dz * float(1u + lz));
        if (ok) {
            a.blo[C] = make_float4(cx.lo, cy.lo, cz.lo, 0.0f);
            a.bhi[C] = make_float4(cx.hi, cy.hi, cz.hi, 0.0f);
        }
        Iv ct;
        if (isnew || pfast || width <= eps) {
            ct = a.swept ? I(time - a.swept_window, time) : I_FROM(time);
            cmeta &= ~kFastBit;
        } else {
            ct = I(cst.y, time);
        }
        float leaf_w = width;
        if (a.point_sample && width <= eps) {
            cx = I_FROM(MID(cx));
            cy = I_FROM(MID(cy));
            cz = I_FROM(MID(cz));
            leaf_w = 0.0f;
        }
        const bool guard = ok && (isnew || cst.x >= 0.0f || (cmeta & kChildMask) == 0);
        const Iv targ = I_FROM_UNORDERED(ct.lo, ct.hi);
        const bool td = ct.lo != ct.hi;

        // ---- program: the group's pruned program when every child's t-window lies inside its
        //      parent's (box containment holds by construction), else the full tape ----
        const FwdIns *fprog = a.fprog;
        const RevIns *rprog = a.rprog;
        uint32_t nprog = a.nfwd;
        if (prune && level > 0) {
            const uint32_t gbase = gdesc[2u * g], gm = gdesc[2u * g + 1u];
            const bool has = active && pidx < a.prune_cap;
            const float wlo = has ? __uint_as_float(prec_in[pidx]) : 0.0f;
            const float whi = has ? __uint_as_float(prec_in[a.prune_cap + pidx]) : 0.0f;
            const bool inside = targ.lo >= wlo && targ.hi <= whi && cx.lo >= plo.x && cx.hi <= phi.x &&
                                cy.lo >= plo.y && cy.hi <= phi.y && cz.lo >= plo.z && cz.hi <= phi.z;
            const bool outside = guard && !inside;
            if (gm != 0u && ballot(outside) == 0) {
                const uint32_t pad = (gm + 3u) / 4u * 4u + 8u;
                fprog = a.pool + gbase;
                rprog = reinterpret_cast<const RevIns *>(a.pool + gbase + pad);
                nprog = gm;
            }
        }

        // ---- tape evaluation (wave-uniform) ----
        Rlo[a.arg_slot[0]] = cx.lo; Rhi[a.arg_slot[0]] = cx.hi;
        Rlo[a.arg_slot[1]] = cy.lo; Rhi[a.arg_slot[1]] = cy.hi;
        Rlo[a.arg_slot[2]] = cz.lo; Rhi[a.arg_slot[2]] = cz.hi;
        Rlo[a.arg_slot[3]] = targ.lo; Rhi[a.arg_slot[3]] = targ.hi;
        Iv field = I_FROM(0.0f);
        ch.word = 0u;
        ch.n = 0u;
        const uint64_t need_eval = ballot(guard);
        if (need_eval) {
            const uint32_t fc = prune ? run_forward<REV, true>(fprog, nprog, ps, Rlo, Rhi, stk, ch)
                                      : run_forward<REV, false>(fprog, nprog, ps, Rlo, Rhi, stk, ch);
            field = I(Rlo[a.out_slot], Rhi[a.out_slot]);
            if (guard) {
                fwd_ops += fc;
                n_pruned += a.nfwd - nprog;
            }
        }
        Iv df = I_FROM(0.0f), df_dt = I_FROM(0.0f);
        float4 gradv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float terr = cst.z;
        if (REV) {
            const uint64_t need_rev = ballot(guard && td);
            if (need_rev) {
                uint32_t ops = run_reverse(rprog, nprog, ps, Rlo, Rhi, stk);
                if (guard && td) {
                    rev_ops += ops;
                    const Iv s0 = I(Rlo[0], Rhi[0]), s1 = I(Rlo[1], Rhi[1]), s2 = I(Rlo[2], Rhi[2]);
                    // the sign of a zero sensitivity depends on which pruned instructions ran;
                    // dx/dy/dz are stored with +0 for any zero
                    gradv = make_float4(MID(s0) + 0.0f, MID(s1) + 0.0f, MID(s2) + 0.0f, 0.0f);   // 3967-3969
                    df_dt = I(Rlo[3], Rhi[3]);                                   // 3972
                    const float dt = DIAMETER(targ);
                    df = MUL_IF(df_dt, dt);
                    const Iv W = MUL_IF(I(-1.0f, 1.0f), DIAMETER(field));
                    df = CAP_II(df, W);
                    if (a.lipschitz && ABS_I(df).hi != 0.0f) {                    // 3980-3992
                        const float lb =
                            ABS_I(SQRT_I(ADD_II(SQR_I(s0), ADD_II(SQR_I(s1), SQR_I(s2))))).lo;
                        if (!is_nan_w(lb)) df = DIV_II(df, I_FROM(lb));
                    }
                    const float inf = __uint_as_float(0x7f800000u);
                    if (df.lo == -inf || df.hi == inf || is_nan_w(df.lo) || is_nan_w(df.hi)) df = I_FROM(a.delta);
                    terr = terr + wmin(wmax(ABS_I(df).hi, 0.0f), a.delta);       // 3996
                }
            }
        }

        // ---- fast flag and classification (compute.wgsl:4003-4045) ----
        float f = cst.x;
        bool enq = false, drop_children = false;
        if (guard) {
            if (!a.swept && a.fastenabled) {
                if (ABS_I(df).hi >= a.delta * a.fast_multiplier) cmeta |= kFastBit;
                else cmeta &= ~kFastBit;
            }
            if (field.lo <= 0.0f && 0.0f <= field.hi) {
                f = 0.0f;
                if (leaf_w > eps) {
                    const float inf = __uint_as_float(0x7f800000u);
                    if (isnew || is_nan_w(df_dt.lo) || is_nan_w(df_dt.hi) || fabsf(df_dt.lo) == inf ||
                        fabsf(df_dt.hi) == inf || pfast || ct.lo == ct.hi || terr >= a.delta ||
                        (cmeta & kChildMask) == 0) {
                        terr = 0.0f;
                        enq = true;
                    }
                }
            } else {
                if (field.hi <= 0.0f) f = -1.0f;
                else if (!a.swept) f = 1.0f;
                if (!a.swept && (cmeta & kChildMask) != 0) drop_children = true;
            }
        }
        const uint32_t old_child = cmeta & kChildMask;
        if (drop_children) cmeta &= ~kChildMask;
        if (ok) a.st[C] = make_float4(f, time, terr, __uint_as_float(cmeta));
        // dx/dy/dz: written by the reverse pass; zeroed when the node is allocated, so a
        // reused node never shows a previous occupant's values.
        if (ok && (isnew || (REV && guard && td))) a.grad[C] = gradv;

        // ---- compaction of the outputs ----
        const bool next_fwd_only =
            !a.swept && ((cmeta & kChildMask) == 0 || (cmeta & kFastBit) || width / 2.0f <= eps);
        const uint32_t at_f = wave_append(enq && next_fwd_only, C, next_f, &cnt->list_count[level + 1][0],
                                          a.list_cap, &cnt->frame_error, kErrQueue);
        const uint32_t at_r = wave_append(enq && !next_fwd_only, C, next_r, &cnt->list_count[level + 1][1],
                                          a.list_cap, &cnt->frame_error, kErrQueue);
        wave_append(drop_children, old_child - 1u, a.dlists[0], &cnt->dlist_count[0], a.dlist_cap,
                    &cnt->frame_error, kErrQueue);

        // ---- liveness of the enqueued cells, for their children's pruned programs ----
        if (prune && ballot(enq) != 0) {
            uint32_t *rec = nullptr;
            if (at_f < a.prune_cap) rec = prec_f + at_f;
            if (at_r < a.prune_cap) rec = prec_r + at_r;
            if (rec) {
                rec[0] = __float_as_uint(targ.lo);
                rec[a.prune_cap] = __float_as_uint(targ.hi);
            }
            live_walk(fprog, nprog, a.out_slot, ch, a.nchunks, rec, a.prune_cap);
        }
        n_cells += guard ? 1 : 0;
        n_rev += (guard && td) ? 1 : 0;
        n_enq += enq ? 1 : 0;
    }
    // ---- statistics: one atomic per wavefront ----
    n_cells = wave_sum64(n_cells);
    n_rev = wave_sum64(n_rev);
    n_par = wave_sum64(n_par);
    n_enq = wave_sum64(n_enq);
    rev_ops = wave_sum64(rev_ops);
    fwd_ops = wave_sum64(fwd_ops);
    n_pruned = wave_sum64(n_pruned);
    if (lane == 0) {
        if (n_cells) atomicAdd((unsigned long long *)&cnt->stats[kStatCells], (unsigned long long)n_cells);
        if (n_rev) atomicAdd((unsigned long long *)&cnt->stats[kStatRevCells], (unsigned long long)n_rev);
        if (n_par) atomicAdd((unsigned long long *)&cnt->stats[kStatParents], (unsigned long long)n_par);
        if (n_enq) atomicAdd((unsigned long long *)&cnt->stats[kStatEnqueued], (unsigned long long)n_enq);
        if (rev_ops) atomicAdd((unsigned long long *)&cnt->stats[kStatRevOps], (unsigned long long)rev_ops);
        if (fwd_ops) atomicAdd((unsigned long long *)&cnt->stats[kStatFwdOps], (unsigned long long)fwd_ops);
        if (n_pruned) atomicAdd((unsigned long long *)&cnt->stats[kStatPruned], (unsigned long long)n_pruned);
    }
}

// Group-program generator: for every 8 consecutive entries of a frontier list (the
// parents one wavefront of the next k_eval will take), the union of their liveness records
// becomes one dense program, forward and reverse, each MIN/MAX record tagged with the choice
// every parent decided.  Groups with an entry lacking a record run the full program.
__global__ void __launch_bounds__(256) k_prune(KArgs a, int level, int which) {
    Counters *cnt = a.cnt;
    const uint32_t count = min(cnt->list_count[level][which], a.list_cap);
    if (count == 0) return;
    const uint32_t lane = lane_id();
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t *prec = a.prune_rec[level & 1][which];
    uint32_t *desc = a.gdesc[which];
    const uint32_t nc = a.nchunks, cap = a.prune_cap;
    const uint32_t ngroups = (count + 7u) / 8u;
    const uint32_t *ffull = reinterpret_cast<const uint32_t *>(a.fprog);
    const uint32_t *rfull = reinterpret_cast<const uint32_t *>(a.rprog);
    for (uint32_t g = wave; g < ngroups; g += nwaves) {
        const uint32_t e0 = g * 8u;
        const uint32_t ne = count - e0 < 8u ? count - e0 : 8u;
        if (e0 + ne > a.prune_cap) {
            if (lane == 0) { desc[2u * g] = 0u; desc[2u * g + 1u] = 0u; }
            continue;
        }
        // union of the entries' live words; lane t holds chunks t and t + 64
        uint32_t u0 = 0, u1 = 0;
        const uint32_t c0 = lane, c1 = lane + 64u;
        for (uint32_t i = 0; i < ne; ++i) {
            const uint32_t *r = prec + size_t(kRecHead) * cap + e0 + i;
            if (c0 < nc) u0 |= r[size_t(c0) * cap];
            if (c1 < nc) u1 |= r[size_t(c1) * cap];
        }
        // exclusive prefix of per-chunk record counts, in chunk order
        const uint32_t n0 = __popc(u0), n1 = __popc(u1);
        uint32_t s0 = n0, s1 = n1;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y0 = uint32_t(__shfl_up(int(s0), off, 64)), y1 = uint32_t(__shfl_up(int(s1), off, 64));
            if (lane >= uint32_t(off)) { s0 += y0; s1 += y1; }
        }
        const uint32_t tot0 = bcast(s0, 63), tot1 = bcast(s1, 63);
        const uint32_t m = tot0 + tot1;
        const uint32_t off0 = s0 - n0, off1 = tot0 + s1 - n1;
        const uint32_t pad = (m + 3u) / 4u * 4u + 8u;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&cnt->pool_used[level], 2u * pad);
        base = bcast(base, 0);
        if (base + 2u * pad > a.pool_cap) {
            if (lane == 0) { desc[2u * g] = 0u; desc[2u * g + 1u] = 0u; }
            continue;
        }
        uint32_t *fw = reinterpret_cast<uint32_t *>(a.pool + base);
        uint32_t *rw = reinterpret_cast<uint32_t *>(a.pool + base + pad);
        for (int half = 0; half < 2; ++half) {
            const uint32_t c = half ? c1 : c0;
            uint32_t u = half ? u1 : u0, at = half ? off1 : off0;
            if (c >= nc) continue;
            uint32_t Lw[8], Rw[8], Vw[8];
#pragma unroll
            for (uint32_t i = 0; i < 8; ++i) {
                const uint32_t *r = prec + size_t(kRecHead + c) * cap + e0 + (i < ne ? i : 0u);
                Vw[i] = i < ne ? r[0] : 0u;
                Lw[i] = i < ne ? r[size_t(nc) * cap] : 0u;
                Rw[i] = i < ne ? r[size_t(2u * nc) * cap] : 0u;
            }
            while (u) {
                const uint32_t b = __builtin_ctz(u);
                u &= u - 1u;
                const uint32_t k = c * 32u + b;
                uint32_t modes = 0;
#pragma unroll
                for (uint32_t i = 0; i < 8; ++i) {
                    const uint32_t live = (Vw[i] >> b) & 1u;
                    const uint32_t md = live ? (((Lw[i] >> b) & 1u) | (((Rw[i] >> b) & 1u) << 1)) : 3u;
                    modes |= md << (2u * i);
                }
                const uint32_t *f = ffull + size_t(k) * 4u;
                uint32_t *fo = fw + size_t(at) * 4u;
                fo[0] = f[0];
                fo[1] = f[1];
                fo[2] = f[2];
                fo[3] = (f[3] & 0xFFFFu) | (modes << 16);
                const uint32_t *rr = rfull + size_t(a.nfwd - 1u - k) * 4u;
                uint32_t *ro = rw + size_t(m - 1u - at) * 4u;
                ro[0] = rr[0];
                ro[1] = rr[1];
                ro[2] = rr[2];
                ro[3] = (rr[3] & 0xFFFFu) | (modes << 16);
                ++at;
            }
        }
        // zero padding records after both programs (the interpreter prefetches past the end)
        for (uint32_t i = m * 4u + lane; i < pad * 4u; i += 64u) {
            fw[i] = 0u;
            rw[i] = 0u;
        }
        if (lane == 0) { desc[2u * g] = base; desc[2u * g + 1u] = m; }
    }
}

// Dealloc cascade step (compute.wgsl:3533-3543): every node of a freed block that has
// children passes its child block on; the block goes to the pending free list.
__global__ void __launch_bounds__(256) k_dealloc(KArgs a, int step) {
    Counters *cnt = a.cnt;
    const uint32_t count = min(cnt->dlist_count[step], a.dlist_cap);
    if (count == 0) return;
    const uint32_t *in = a.dlists[step & 1];
    uint32_t *out = a.dlists[(step + 1) & 1];
    const uint32_t lane = lane_id();
    const uint32_t o = lane & 7u;
    // wave index made visibly uniform: the grid-stride loop and everything inside it then
    // compile as scalar control flow
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    uint64_t freed = 0;
    for (uint32_t g = wave; g * 8u < count; g += nwaves) {
        const uint32_t idx = g * 8u + (lane >> 3);
        const bool active = idx < count;
        const uint32_t b = active ? in[idx] : 0u;
        const uint32_t m = active ? meta_of(a.st[b * 8u + o]) & kChildMask : 0u;
        wave_append(m != 0, m - 1u, out, &cnt->dlist_count[step + 1], a.dlist_cap, &cnt->frame_error, kErrQueue);
        wave_append(active && o == 0, b, a.pending_free, &cnt->pending, a.nblocks, &cnt->frame_error, kErrQueue);
        freed += (active && o == 0) ? 1 : 0;
    }
    freed = wave_sum64(freed);
    if (lane == 0 && freed) atomicAdd((unsigned long long *)&cnt->stats[kStatDeallocs], (unsigned long long)freed);
}

// Free-list merge: blocks handed out from the top of the stack are gone, freed blocks go on top.
__global__ void __launch_bounds__(256) k_frame_end(KArgs a) {
    Counters *cnt = a.cnt;
    const uint32_t T = cnt->alloc_tickets, F = cnt->free_top, P = min(cnt->pending, a.nblocks);
    const uint32_t used = T < F ? T : F;
    const uint32_t base = F - used;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P && base + i < a.nblocks; i += gridDim.x * blockDim.x)
        a.free_stack[base + i] = a.pending_free[i];
}

__global__ void k_frame_commit(KArgs a) {
    Counters *cnt = a.cnt;
    const uint32_t T = cnt->alloc_tickets, F = cnt->free_top, P = min(cnt->pending, a.nblocks);
    const uint32_t used = T < F ? T : F;
    uint32_t bump = cnt->bump + (T > F ? T - F : 0u);
    cnt->bump = bump < a.nblocks - 1u ? bump : a.nblocks - 1u;
    cnt->free_top = min(F - used + P, a.nblocks);
    cnt->stats[kStatLive] = uint64_t(cnt->bump) - uint64_t(cnt->free_top);
    cnt->error |= cnt->frame_error;
}

__global__ void k_clear_nodes(KArgs a, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        a.st[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

__global__ void k_reset(KArgs a, float x0, float y0, float z0, float x1, float y1, float z1) {
    Counters *c = a.cnt;
    a.st[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);   // main.rs:749-763: f=0, t=0, no children
    a.blo[0] = make_float4(x0, y0, z0, 0.0f);
    a.bhi[0] = make_float4(x1, y1, z1, 0.0f);
    a.grad[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    c->root_pending = 1;
    c->alloc_tickets = 0;
    c->free_top = 0;
    c->bump = 0;
    c->pending = 0;
    c->error = 0;
    c->frame_error = 0;
}

__global__ void k_enqueue_root(KArgs a) { a.cnt->root_pending = 1; }

// ---------------------------------------------------------------- trace (rt_main)

struct V3 { float x, y, z; };
TSI_DI V3 v3(float x, float y, float z) { return V3{x, y, z}; }
TSI_DI float vlength(V3 a) { return sqrtf((a.x * a.x + a.y * a.y) + a.z * a.z); }
TSI_DI float wsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

constexpr float kTrEps = 0.0001f;     // compute.wgsl:4140
constexpr float kDarken = 0.98f;      // compute.wgsl:3507

// compute.wgsl:4059-4070
TSI_DI V3 efficient_slabs(V3 p0, V3 p1, V3 o, V3 inv) {
    const float t0x = (p0.x - o.x) * inv.x, t0y = (p0.y - o.y) * inv.y, t0z = (p0.z - o.z) * inv.z;
    const float t1x = (p1.x - o.x) * inv.x, t1y = (p1.y - o.y) * inv.y, t1z = (p1.z - o.z) * inv.z;
    const float d1 = wmax(wmin(t0x, t1x), wmax(wmin(t0y, t1y), wmin(t0z, t1z)));
    const float d2 = wmin(wmax(t0x, t1x), wmin(wmax(t0y, t1y), wmax(t0z, t1z)));
    return v3((d1 <= d2 && d2 >= 0.0f) ? 1.0f : 0.0f, wmax(d1, 0.0f), wmax(d2, 0.0f));
}

TSI_DI float mylerp(float a, float b, float t) {   // compute.wgsl:4116
    const float tt = wmin(wmax(t, 0.0f), 1.0f);
    return a * (1.0f - tt) + b * tt;
}

// compute.wgsl:4189-4403 for the non-cinematic, non-normal_viz colour; the stack lives in LDS.
TSI_DI void trace_ray(const KArgs &a, uint32_t *stk, V3 ro, V3 rd, float time, V3 &col, float &dist,
                      uint32_t &visits) {
    constexpr uint32_t FAIL = 0xFFFFFFFFu;
    uint32_t sp = 0;
    V3 c = v3(1.0f, 1.0f, 1.0f);
    V3 o2 = ro;
    bool intersected = false;
    const V3 inv = v3(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
    auto push = [&](uint32_t v) { if (sp != 16u) { stk[sp * 256u] = v; ++sp; } };
    auto peak = [&]() -> uint32_t { return sp == 0u ? FAIL : stk[(sp - 1u) * 256u]; };
    auto box = [&](uint32_t n, V3 &p0, V3 &p1) {
        const float4 lo = a.blo[n], hi = a.bhi[n];
        p0 = v3(lo.x, lo.y, lo.z);
        p1 = v3(hi.x, hi.y, hi.z);
    };
    push(0u);
    uint32_t region = 0u;
    V3 p0, p1;
    box(region, p0, p1);
    V3 hit = efficient_slabs(p0, p1, o2, inv);
    if (hit.x == 0.0f) { col = c; dist = -1.0f; return; }
    float sms = 0.0f;
    for (uint32_t iter = 0; iter < 5000u; ++iter) {
        if (intersected) break;
        if (sp == 0u) { col = c; dist = -1.0f; return; }
        --sp;
        region = stk[sp * 256u];
        if (hit.x == 0.0f) {
            const uint32_t top = peak();
            if (top != FAIL) {
                box(top, p0, p1);
                hit = efficient_slabs(p0, p1, o2, inv);
            }
            continue;
        }
        const float ex = p1.x - p0.x, ey = p1.y - p0.y, ez = p1.z - p0.z;
        const float rox = ((o2.x + (hit.y + kTrEps) * rd.x) - p0.x) * 2.0f / ex;
        const float roy = ((o2.y + (hit.y + kTrEps) * rd.y) - p0.y) * 2.0f / ey;
        const float roz = ((o2.z + (hit.y + kTrEps) * rd.z) - p0.z) * 2.0f / ez;
        const float rdx = rd.x * (2.0f / ex), rdy = rd.y * (2.0f / ey), rdz = rd.z * (2.0f / ez);
        int32_t mx = f2i_sat(floorf(rox)), my = f2i_sat(floorf(roy)), mz = f2i_sat(floorf(roz));
        const float ddx = 1.0f / fabsf(rdx), ddy = 1.0f / fabsf(rdy), ddz = 1.0f / fabsf(rdz);
        const float sgx = wsign(rdx), sgy = wsign(rdy), sgz = wsign(rdz);
        const int32_t stx = f2i_sat(sgx), sty = f2i_sat(sgy), stz = f2i_sat(sgz);
        float sdx = ((sgx * (float(mx) - rox)) + sgx * 0.5f + 0.5f) * ddx;
        float sdy = ((sgy * (float(my) - roy)) + sgy * 0.5f + 0.5f) * ddy;
        float sdz = ((sgz * (float(mz) - roz)) + sgz * 0.5f + 0.5f) * ddz;
        float mkx = 0.0f, mky = 0.0f, mkz = 0.0f;
        const V3 old = o2;
        const uint32_t rmeta = meta_of(a.st[region]) & kChildMask;
        for (int i = 0; i < 32; ++i) {
            if (wmax(p1.x - p0.x, wmax(p1.y - p0.y, p1.z - p0.z)) < 1.0f) {
                c.x = c.x * kDarken; c.y = c.y * kDarken; c.z = c.z * kDarken;
            }
            const float t = ((vlength(v3(mkx * (sdx - ddx), mky * (sdy - ddy), mkz * (sdz - ddz))) + hit.y) + kTrEps) + sms;
            o2 = v3(old.x + t * rd.x, old.y + t * rd.y, old.z + t * rd.z);
            sms = 0.0f;
            if (t >= hit.z - kTrEps) {
                o2 = v3(old.x + hit.z * rd.x, old.y + hit.z * rd.y, old.z + hit.z * rd.z);
                const uint32_t top = peak();
                if (top != FAIL) {
                    box(top, p0, p1);
                    hit = efficient_slabs(p0, p1, o2, inv);
                }
                break;
            }
            const int32_t ci = int32_t(uint32_t(mx) + 2u * uint32_t(my) + 4u * uint32_t(mz));
            if (ci < 0 || ci >= 8) break;
            if (rmeta != 0u) {
                const uint32_t child = (rmeta - 1u) * 8u + uint32_t(ci);
                const float4 cs = a.st[child];
                ++visits;
                if (cs.x <= 0.0f) {
                    const float4 lo = a.blo[child], hi = a.bhi[child];
                    const V3 tp0 = v3(lo.x - kTrEps, lo.y - kTrEps, lo.z - kTrEps);
                    const V3 tp1 = v3(hi.x + kTrEps, hi.y + kTrEps, hi.z + kTrEps);
                    const V3 th = efficient_slabs(tp0, tp1, o2, inv);
                    if (th.x == 1.0f) {
                        if ((meta_of(cs) & kChildMask) != 0u) {
                            push(region);
                            push(child);
                            p0 = tp0;
                            p1 = tp1;
                            hit = th;
                            break;
                        } else {
                            intersected = true;
                            const float tt = (time - cs.y) * 60.0f;
                            c = v3(c.x - mylerp(0.0f, 0.0f, tt), c.y - mylerp(1.0f, 0.0f, tt), c.z - mylerp(1.0f, 0.0f, tt));
                            break;
                        }
                    }
                }
            }
            const bool bx = sdx <= wmin(sdy, sdz), by = sdy <= wmin(sdz, sdx), bz = sdz <= wmin(sdx, sdy);
            mkx = bx ? 1.0f : 0.0f;
            mky = by ? 1.0f : 0.0f;
            mkz = bz ? 1.0f : 0.0f;
            sdx = sdx + mkx * ddx;
            sdy = sdy + mky * ddy;
            sdz = sdz + mkz * ddz;
            mx = int32_t(uint32_t(mx) + uint32_t(bx ? stx : 0));
            my = int32_t(uint32_t(my) + uint32_t(by ? sty : 0));
            mz = int32_t(uint32_t(mz) + uint32_t(bz ? stz : 0));
        }
    }
    c = v3(wmax(wmin(c.x, 1.0f), 0.0f), wmax(wmin(c.y, 1.0f), 0.0f), wmax(wmin(c.z, 1.0f), 0.0f));
    float d = fabsf(vlength(v3(o2.x - ro.x, o2.y - ro.y, o2.z - ro.z)));
    if (!intersected) d = -1.0f;
    col = c;
    dist = d;
}

__global__ void __launch_bounds__(256) k_trace(KArgs a, TraceArgs t) {
    __shared__ uint32_t stack_lds[16 * 256];
    const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u);
    const uint32_t y = blockIdx.y * 16u + (threadIdx.x >> 4);
    uint32_t visits = 0;
    if (x < t.width && y < t.height) {
        // ray_direction (compute.wgsl:4590-4594) and view * vec4(d, 0) (4630)
        const float xyx = (float(x) - float(t.width) / 2.0f) + 0.5f;
        const float xyy = (float(y) - float(t.height) / 2.0f) + 0.5f;
        V3 d = v3(xyx, -xyy, -t.zdist);
        const float len = vlength(d);
        d = v3(d.x / len, d.y / len, d.z / len);
        const float *m = t.view;
        const V3 g = v3(((m[0] * d.x + m[4] * d.y) + m[8] * d.z) + m[12] * 0.0f,
                        ((m[1] * d.x + m[5] * d.y) + m[9] * d.z) + m[13] * 0.0f,
                        ((m[2] * d.x + m[6] * d.y) + m[10] * d.z) + m[14] * 0.0f);
        const V3 eye = v3(m[12], m[13], m[14]);
        V3 c;
        float dist;
        trace_ray(a, stack_lds + threadIdx.x, eye, g, t.time, c, dist, visits);
        const size_t pi = size_t(y) * t.width + x;
        const float r = (0.0f + c.x) / 1.0f, gg = (0.0f + c.y) / 1.0f, b = (0.0f + c.z) / 1.0f;
        t.rgba[pi] = make_float4(det_pow045(r), det_pow045(gg), det_pow045(b), (0.0f + 1.0f) / 1.0f);
        t.depth[pi] = dist;
    }
    uint64_t v = wave_sum64(visits);
    if (lane_id() == 0 && v) atomicAdd((unsigned long long *)t.visits, (unsigned long long)v);
}

// ---------------------------------------------------------------- launchers

int launch_frame(const KArgs &a, const FrameLaunch &fl, hipStream_t s, hipEvent_t e_eval0, hipEvent_t e_eval1) {
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(64), 0, s, a);
    if (e_eval0) hipEventRecord(e_eval0, s);
    for (int d = 0; d < fl.levels; ++d) {
        if (a.prune && d > 0) {
            hipLaunchKernelGGL(k_prune, dim3(fl.grid_misc), dim3(256), 0, s, a, d, 0);
            hipLaunchKernelGGL(k_prune, dim3(fl.grid_misc), dim3(256), 0, s, a, d, 1);
        }
        hipLaunchKernelGGL(k_eval<false>, dim3(fl.grid_fwd), dim3(256), 0, s, a, d, fl.time);
        hipLaunchKernelGGL(k_eval<true>, dim3(fl.grid_rev), dim3(256), 0, s, a, d, fl.time);
    }
    if (e_eval1) hipEventRecord(e_eval1, s);
    for (int c = 0; c <= fl.levels; ++c)
        hipLaunchKernelGGL(k_dealloc, dim3(fl.grid_misc), dim3(256), 0, s, a, c);
    hipLaunchKernelGGL(k_frame_end, dim3(fl.grid_misc), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_frame_commit, dim3(1), dim3(1), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_reset(const KArgs &a, const float box[6], uint64_t nodes, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_nodes, dim3(1024), dim3(256), 0, s, a, nodes);
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, s, a, box[0], box[1], box[2], box[3], box[4], box[5]);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_enqueue_root(const KArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_enqueue_root, dim3(1), dim3(1), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_trace(const KArgs &a, const TraceArgs &t, hipStream_t s) {
    dim3 grid((t.width + 15u) / 16u, (t.height + 15u) / 16u);
    hipLaunchKernelGGL(k_trace, grid, dim3(256), 0, s, a, t);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tsi
