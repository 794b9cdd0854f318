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

// Appends one element per lane where p holds to list[0 .. cap), count at *cnt.
// One atomic per wavefront; returns false (and flags an error) on overflow.
TSI_DI void wave_append(bool p, uint32_t value, uint32_t *list, uint32_t *cnt, uint32_t cap, uint32_t *err,
                        uint32_t errbit) {
    uint64_t m = ballot(p);
    if (m == 0) return;
    uint32_t n = popc64(m), base = 0;
    if (lane_id() == 0) base = atomicAdd(cnt, n);
    base = bcast(base, 0);
    if (p) {
        uint32_t at = base + rank_in(m);
        if (at < cap) list[at] = value;
        else atomicOr(err, errbit);
    }
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

// One forward step (compute.wgsl:3642-3772): record words w0 = {form, a, b, d},
// w1 = {push, npush}, w2 = literal.
template <bool PUSH>
TSI_DI void fwd_step(uint32_t w0, uint32_t w1, uint32_t w2, Slots &Rlo, Slots &Rhi, float2 *__restrict__ stk) {
    const uint32_t form = w0 & 0xFFu, ia = (w0 >> 8) & 0xFFu, ib = (w0 >> 16) & 0xFFu, id = w0 >> 24;
    const Iv l = I(Rlo[ia], Rhi[ia]);
    const Iv r = I(Rlo[ib], Rhi[ib]);
    const float c = __uint_as_float(w2);
    Iv v;
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
    case F_MIN_RR: v = MIN_II(l, r); break;
    case F_MIN_RL: v = MIN_IF(l, c); break;
    case F_MAX_RR: v = MAX_II(l, r); break;
    case F_MAX_RL: v = MAX_IF(l, c); break;
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
    if (PUSH) {
        const uint32_t npush = (w1 >> 16) & 0xFFu;
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
// Programs are padded with zero records to a multiple of four plus one batch.
constexpr uint32_t kBatch = 4;

// Forward pass over the lowered tape (compute.wgsl:3633-3772).  When PUSH, the operand
// values the reverse rules need are saved to this wave's stack region.
template <bool PUSH>
TSI_DI void run_forward(const FwdIns *prog, uint32_t nfwd, Slots &Rlo, Slots &Rhi, float2 *__restrict__ stk) {
    CWords base = (CWords)prog;
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
            if (k + i < nfwd) fwd_step<PUSH>(cur[4 * i], cur[4 * i + 1], cur[4 * i + 2], Rlo, Rhi, stk);
#pragma unroll
        for (int i = 0; i < int(4 * kBatch); ++i) cur[i] = nxt[i];
    }
}

// One reverse step (compute.wgsl:3786-3963): w0 = {form, sj, sa, sb}, w1 = {flags, side, pop},
// w2 = literal, w3 = {cost, npush}; ops_in = the saved operand values (left.lo, left.hi, right.lo,
// right.hi; a single saved interval sits in .xy).  Returns the op count when not short-circuited.
TSI_DI uint32_t rev_step(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, Slots &Slo, Slots &Shi,
                         float4 ops_in) {
    const Iv Z = I_FROM(0.0f), ONE = I_FROM(1.0f), NINF = I_FROM(__uint_as_float(0xff800000u));
    const uint32_t form = w0 & 0xFFu, isj = (w0 >> 8) & 0xFFu, isa = (w0 >> 16) & 0xFFu, isb = w0 >> 24;
    const uint32_t flags = w1 & 0xFFu, side = (w1 >> 8) & 0xFFu, pop = w1 >> 16;
    const float c = __uint_as_float(w2);
    const uint32_t cost = w3 & 0xFFFFu;
    Iv sj = (flags & R_SJ_ONE) ? ONE : Z;
    if (!(flags & (R_SJ_ONE | R_SJ_ZERO))) sj = I(Slo[isj], Shi[isj]);
    const bool sj0 = sj.lo == 0.0f && sj.hi == 0.0f;
    uint32_t ops = 0;
    (void)pop;
    if (side == 1) {
        const Iv l = I(ops_in.x, ops_in.y);
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
        case F_MIN_RL: x = sel(l.hi < c, ONE, sel(c < l.lo, Z, ONE)); break;
        case F_MAX_RL: x = sel(l.hi < c, Z, ONE); break;
        default: x = NINF; break;
        }
        const Iv old = (flags & R_FT_A) ? Z : I(Slo[isa], Shi[isa]);
        Iv nv = ADD_II(MUL_II(fix(x), sj), old);
        const bool sc = sj0 && (flags & R_U1_A);
        nv = sel(sc, Z, nv);
        ops = sc ? 0u : cost;
        Slo[isa] = nv.lo;
        Shi[isa] = nv.hi;
    } else if (side == 0) {
        const Iv l = I(ops_in.x, ops_in.y), r = I(ops_in.z, ops_in.w);
        Iv x0, x1;
        switch (form) {
        case F_ADD_RR: x0 = ONE; x1 = ONE; break;
        case F_SUB_RR: x0 = ONE; x1 = I_FROM(-1.0f); break;
        case F_MUL_RR: x0 = r; x1 = l; break;
        case F_MIN_RR: {   // MIN_II_RR (compute.wgsl:2884)
            const bool lt = l.hi < r.lo, gt = r.hi < l.lo;
            x0 = sel(lt, ONE, sel(gt, Z, ONE));
            x1 = sel(lt, Z, ONE);
            break;
        }
        case F_MAX_RR: {   // MAX_II_RR (compute.wgsl:2894)
            const bool lt = l.hi < r.lo, gt = r.hi < l.lo;
            x0 = sel(lt, Z, ONE);
            x1 = sel(lt, ONE, sel(gt, Z, ONE));
            break;
        }
        case F_DIV_RR: x0 = DIV_II(ONE, r); x1 = NEG_I(DIV_II(l, SQR_I(r))); break;
        case F_POW_RR: case F_ATAN2_RR: rare_reverse(form, l, r, c, &x0, &x1); break;
        case F_NOP_RR: x0 = Z; x1 = Z; break;
        default: x0 = NINF; x1 = NINF; break;
        }
        const Iv olda = (flags & R_FT_A) ? Z : I(Slo[isa], Shi[isa]);
        const Iv oldb = (flags & R_FT_B) ? Z : I(Slo[isb], Shi[isb]);
        Iv na = ADD_II(MUL_II(fix(x0), sj), olda);
        Iv nb = ADD_II(MUL_II(fix(x1), sj), oldb);
        const bool sc = sj0 && (flags & R_U1_A) && (flags & R_U1_B);
        na = sel(sc, Z, na);
        nb = sel(sc, Z, nb);
        ops = sc ? 0u : cost;
        Slo[isa] = na.lo;
        Shi[isa] = na.hi;
        Slo[isb] = nb.lo;
        Shi[isb] = nb.hi;
    } else if (side == 2) {
        const Iv r = I(ops_in.x, ops_in.y);
        Iv x;
        switch (form) {
        case F_SUB_LR: x = I_FROM(-1.0f); break;
        case F_DIV_LR: x = NEG_I(DIV_II(I_FROM(c), SQR_I(r))); break;
        default: x = NINF; break;
        }
        const Iv old = (flags & R_FT_B) ? Z : I(Slo[isb], Shi[isb]);
        Iv nv = ADD_II(MUL_II(fix(x), sj), old);
        const bool sc = sj0 && (flags & R_U1_B);
        nv = sel(sc, Z, nv);
        ops = sc ? 0u : cost;
        Slo[isb] = nv.lo;
        Shi[isb] = nv.hi;
    }
    return ops;
}

// Saved operands of one reverse record: npush (w3 bits 16-23) intervals at position pop.
TSI_DI float4 load_saved(uint32_t w1, uint32_t w3, const float2 *__restrict__ stk) {
    const uint32_t npush = (w3 >> 16) & 0xFFu;
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
TSI_DI uint32_t run_reverse(const RevIns *prog, uint32_t nrev, Slots &Slo, Slots &Shi,
                            const float2 *__restrict__ stk) {
    uint32_t ops = 0;
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
            if (k + i < nrev) ops += rev_step(cur[4 * i], cur[4 * i + 1], cur[4 * i + 2], cur[4 * i + 3], Slo, Shi, scur[i]);
#pragma unroll
        for (int i = 0; i < int(4 * kBatch); ++i) { cur[i] = nxt[i]; nxt[i] = nn[i]; }
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) scur[i] = snxt[i];
    }
    return ops;
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
    const uint32_t count = cnt->list_count[level][which];
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

    uint64_t n_cells = 0, n_rev = 0, n_par = 0, n_enq = 0, rev_ops = 0;
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

        // ---- tape evaluation (wave-uniform) ----
        Rlo[a.arg_slot[0]] = cx.lo; Rhi[a.arg_slot[0]] = cx.hi;
        Rlo[a.arg_slot[1]] = cy.lo; Rhi[a.arg_slot[1]] = cy.hi;
        Rlo[a.arg_slot[2]] = cz.lo; Rhi[a.arg_slot[2]] = cz.hi;
        Rlo[a.arg_slot[3]] = targ.lo; Rhi[a.arg_slot[3]] = targ.hi;
        Iv field;
        const uint64_t need_eval = ballot(guard);
        if (need_eval) {
            run_forward<REV>(a.fprog, a.nfwd, Rlo, Rhi, stk);
            field = I(Rlo[a.out_slot], Rhi[a.out_slot]);
        } else {
            field = I_FROM(0.0f);
        }
        Iv df = I_FROM(0.0f), df_dt = I_FROM(0.0f);
        float4 gradv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float terr = cst.z;
        if (REV) {
            const uint64_t need_rev = ballot(guard && td);
            if (need_rev) {
                uint32_t ops = run_reverse(a.rprog, a.nrev, Rlo, Rhi, stk);
                if (guard && td) {
                    rev_ops += ops;
                    const Iv s0 = I(Rlo[0], Rhi[0]), s1 = I(Rlo[1], Rhi[1]), s2 = I(Rlo[2], Rhi[2]);
                    gradv = make_float4(MID(s0), MID(s1), MID(s2), 0.0f);         // 3967-3969
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
        wave_append(enq && next_fwd_only, C, next_f, &cnt->list_count[level + 1][0], a.list_cap, &cnt->frame_error,
                    kErrQueue);
        wave_append(enq && !next_fwd_only, C, next_r, &cnt->list_count[level + 1][1], a.list_cap,
                    &cnt->frame_error, kErrQueue);
        wave_append(drop_children, old_child - 1u, a.dlists[0], &cnt->dlist_count[0], a.dlist_cap,
                    &cnt->frame_error, kErrQueue);
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
    if (lane == 0) {
        if (n_cells) atomicAdd((unsigned long long *)&cnt->stats[kStatCells], (unsigned long long)n_cells);
        if (n_rev) atomicAdd((unsigned long long *)&cnt->stats[kStatRevCells], (unsigned long long)n_rev);
        if (n_par) atomicAdd((unsigned long long *)&cnt->stats[kStatParents], (unsigned long long)n_par);
        if (n_enq) atomicAdd((unsigned long long *)&cnt->stats[kStatEnqueued], (unsigned long long)n_enq);
        if (rev_ops) atomicAdd((unsigned long long *)&cnt->stats[kStatRevOps], (unsigned long long)rev_ops);
    }
}

// Dealloc cascade step (compute.wgsl:3533-3543): every node of a freed block that has
// children passes its child block on; the block goes to the pending free list.
__global__ void __launch_bounds__(256) k_dealloc(KArgs a, int step) {
    Counters *cnt = a.cnt;
    const uint32_t count = cnt->dlist_count[step];
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
    const uint32_t T = cnt->alloc_tickets, F = cnt->free_top, P = cnt->pending;
    const uint32_t used = T < F ? T : F;
    const uint32_t base = F - used;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x)
        a.free_stack[base + i] = a.pending_free[i];
}

__global__ void k_frame_commit(KArgs a) {
    Counters *cnt = a.cnt;
    const uint32_t T = cnt->alloc_tickets, F = cnt->free_top, P = cnt->pending;
    const uint32_t used = T < F ? T : F;
    uint32_t bump = cnt->bump + (T > F ? T - F : 0u);
    cnt->bump = bump < a.nblocks - 1u ? bump : a.nblocks - 1u;
    cnt->free_top = F - used + P;
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
