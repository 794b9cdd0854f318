// tsi_interp.h -- the tape interpreter: one wavefront runs a lowered program (records in
// constant memory, fetched with scalar loads) for its 64 cells, with the live values of each
// lane in a wave-uniformly indexed register file, and, for the reverse pass, the operand
// values the reverse rules need in a per-wavefront stack.  Programs are either the full tape
// or a group's pruned program (only the records live for one of the wave's 8 parents, with
// MIN/MAX choices the parents decided turned into copies).
#pragma once
#include "tsi_eval.h"
#include "tsi_rules.h"

namespace tsi {

// The interpreter's register file (Slots, tsi_eval.h): kSlots intervals per lane, indexed only
// with wave-uniform slot numbers so the compiler keeps it in VGPRs (s_set_gpr_idx moves).

// Programs are read through the constant address space: the kernel never writes them, so the
// loads become scalar (s_load) and the slot numbers land in SGPRs (no per-lane waterfall).
typedef const __attribute__((address_space(4))) uint32_t *CWords;

// ---- pruned programs -------------------------------------------------------------------
// A program record's word 3 carries, in bits 16-31, the MIN/MAX choice each of the wave's 8
// parents decided (2 bits per parent: 1 = left operand only, 2 = right operand only,
// 3 = undecided).  Full programs carry 0xFFFF.  A parent's decision holds in its children
// (child box and t-window inside the parent's), so a decided MIN/MAX is a copy of its operand.
TSI_DI uint32_t lane_mode(uint32_t w3, uint32_t ps) { return (w3 >> (16u + 2u * ps)) & 3u; }

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
    default: v = fwd_form(form, l, r, c); break;
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
            float2 *p = stk + size_t(w1 & 0xFFFFu) * 64;   // uniform row address
            if (npush == 3) {
                p[lane_id()] = make_float2(r.lo, r.hi);
            } else {
                p[lane_id()] = make_float2(l.lo, l.hi);
                if (npush == 2) p[64 + lane_id()] = make_float2(r.lo, r.hi);
            }
        }
    }
}

// Programs are fetched in 64-byte batches of four 16-byte records, one batch ahead of the
// batch being executed, so the scalar-load latency overlaps four interpreted instructions.
// Programs are padded with zero records (program_stride in tsi_tape.h): the reverse pass
// fetches up to four batches past the start of its last pair of batches.
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
    const Iv Z = I_FROM(0.0f), ONE = I_FROM(1.0f);
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
    const Iv x = I(ops_in.x, ops_in.y), y = I(ops_in.z, ops_in.w);
    uint32_t ops = 0;
    if (side == 1) {
        uint32_t take = 1u;   // the operand receives a contribution
        Iv d;
        switch (form) {
        case F_MIN_RL: {   // MIN_II_RR(l, [c, c]), or the parent's decided choice
            const uint32_t m = lane_mode(w3, ps);
            d = sel(m == 1u, ONE, sel(x.hi < c, ONE, sel(c < x.lo, Z, ONE)));
            take = m & 1u;
            break;
        }
        case F_MAX_RL: {
            const uint32_t m = lane_mode(w3, ps);
            d = sel(m == 1u, ONE, sel(x.hi < c, Z, ONE));
            take = m & 1u;
            break;
        }
        default: d = rev_form(form, x, Z, c).a; break;
        }
        const Iv old = fta ? Z : I(Slo[isa], Shi[isa]);
        Iv nv = take ? ADD_II(MUL_II(fix(d), sj), old) : old;
        const bool sc = sj0 && (flags & R_U1_A);
        nv = sel(sc, Z, nv);
        ops = sc ? 0u : cost;
        Slo[isa] = nv.lo;
        Shi[isa] = nv.hi;
        init |= 1u << isa;
    } else if (side == 0) {
        uint32_t takea = 1u, takeb = 1u;
        Iv2 d;
        switch (form) {
        case F_MIN_RR: {   // MIN_II_RR (compute.wgsl:2884), or the parent's decided choice
            const uint32_t m = lane_mode(w3, ps);
            const bool lt = x.hi < y.lo, gt = y.hi < x.lo;
            d.a = sel(m != 3u, ONE, sel(lt, ONE, sel(gt, Z, ONE)));
            d.b = sel(m != 3u, ONE, sel(lt, Z, ONE));
            takea = m & 1u;
            takeb = m >> 1;
            break;
        }
        case F_MAX_RR: {   // MAX_II_RR (compute.wgsl:2894), or the parent's decided choice
            const uint32_t m = lane_mode(w3, ps);
            const bool lt = x.hi < y.lo, gt = y.hi < x.lo;
            d.a = sel(m != 3u, ONE, sel(lt, Z, ONE));
            d.b = sel(m != 3u, ONE, sel(lt, ONE, sel(gt, Z, ONE)));
            takea = m & 1u;
            takeb = m >> 1;
            break;
        }
        default: d = rev_form(form, x, y, c); break;
        }
        // both previous values are read before either is written (an operand used twice ends
        // with the right-hand contribution, as compute.wgsl:3860-3870 does)
        const Iv olda = fta ? Z : I(Slo[isa], Shi[isa]);
        const Iv oldb = ftb ? Z : I(Slo[isb], Shi[isb]);
        Iv na = takea ? ADD_II(MUL_II(fix(d.a), sj), olda) : olda;
        Iv nb = takeb ? ADD_II(MUL_II(fix(d.b), sj), oldb) : oldb;
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
        const Iv d = rev_form(form, x, Z, c).a;
        const Iv old = ftb ? Z : I(Slo[isb], Shi[isb]);
        Iv nv = ADD_II(MUL_II(fix(d), sj), old);
        const bool sc = sj0 && (flags & R_U1_B);
        nv = sel(sc, Z, nv);
        ops = sc ? 0u : cost;
        Slo[isb] = nv.lo;
        Shi[isb] = nv.hi;
        init |= 1u << isb;
    }
    return ops;
}

// Saved operands of one reverse record: up to two intervals at stack position pop (w1 bits
// 16-31).  Both are always loaded -- the stack has one spare row -- so that every record issues
// the same number of loads and the wait for a batch's operands is a fixed count.
TSI_DI float4 load_saved(uint32_t w1, const float2 *__restrict__ stk) {
    const float2 *p = stk + size_t(w1 >> 16) * 64;   // uniform row address
    const float2 q = p[lane_id()], w = p[64 + lane_id()];
    return make_float4(q.x, q.y, w.x, w.y);
}

// Reverse pass (compute.wgsl:3780-3964) with sensitivities in the register file; returns the
// algorithmic op count of the non-short-circuited steps of this lane.  Batches alternate
// between two register sets A and B: the saved operands of the next batch are loaded while
// the current one executes, into registers the current batch does not use, so no loaded
// value is ever copied (a copy would wait for the load).
TSI_DI uint32_t run_reverse(const RevIns *prog, uint32_t nrev, uint32_t ps, Slots &Slo, Slots &Shi,
                            const float2 *__restrict__ stk) {
    uint32_t ops = 0, init = 0xFu;   // the argument sens slots 0..3 start at [+0, +0]
#pragma unroll
    for (int k = 0; k < 4; ++k) { Slo[k] = 0.0f; Shi[k] = 0.0f; }
    CWords base = (CWords)prog;
    uint32_t wa[4 * kBatch], wb[4 * kBatch];
    float4 sa[kBatch], sb[kBatch];
#pragma unroll
    for (int i = 0; i < int(4 * kBatch); ++i) wa[i] = base[i];
#pragma unroll
    for (int i = 0; i < int(4 * kBatch); ++i) wb[i] = base[4 * kBatch + i];
#pragma unroll
    for (uint32_t i = 0; i < kBatch; ++i) sa[i] = load_saved(wa[4 * i + 1], stk);
    for (uint32_t k = 0; k < nrev; k += 2 * kBatch) {
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) sb[i] = load_saved(wb[4 * i + 1], stk);
        uint32_t wn[4 * kBatch];
        CWords nb = base + size_t(k + 2 * kBatch) * 4u;
#pragma unroll
        for (int i = 0; i < int(4 * kBatch); ++i) wn[i] = nb[i];
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i)
            if (k + i < nrev)
                ops += rev_step(wa[4 * i], wa[4 * i + 1], wa[4 * i + 2], wa[4 * i + 3], ps, Slo, Shi, sa[i], init);
        if (k + kBatch >= nrev) break;
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i) sa[i] = load_saved(wn[4 * i + 1], stk);
        CWords nb2 = base + size_t(k + 3 * kBatch) * 4u;
        uint32_t wn2[4 * kBatch];
#pragma unroll
        for (int i = 0; i < int(4 * kBatch); ++i) wn2[i] = nb2[i];
#pragma unroll
        for (uint32_t i = 0; i < kBatch; ++i)
            if (k + kBatch + i < nrev)
                ops += rev_step(wb[4 * i], wb[4 * i + 1], wb[4 * i + 2], wb[4 * i + 3], ps, Slo, Shi, sb[i], init);
#pragma unroll
        for (int i = 0; i < int(4 * kBatch); ++i) { wa[i] = wn[i]; wb[i] = wn2[i]; }
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

// ---- the interpreter as an evaluation engine of eval_level --------------------------------
struct InterpEngine {
    static constexpr bool kPrune = true;
    const FwdIns *fprog;
    const RevIns *rprog;
    uint32_t n;          // records of the current program
    uint32_t ps;         // parent slot of this lane
    bool choices;        // record MIN/MAX choices (for liveness records)
    Choices ch;

    TSI_DI InterpEngine(const KArgs &a, uint32_t wave, uint32_t lane)
        : fprog(a.fprog), rprog(a.rprog), n(a.nfwd), ps(lane >> 3), choices(a.prune != 0) {
        ch.buf = a.choice + size_t(wave) * a.choice_len * 64 + lane;
        ch.word = 0u;
        ch.n = 0u;
    }

    TSI_DI void select(const KArgs &a, int which, uint32_t g, bool pruned_ok, const uint32_t *, uint32_t, bool) {
        fprog = a.fprog;
        rprog = a.rprog;
        n = a.nfwd;
        if (pruned_ok) {
            const __attribute__((address_space(4))) uint32_t *gdesc =
                (const __attribute__((address_space(4))) uint32_t *)a.gdesc[which];
            const uint32_t gbase = gdesc[2u * g], gm = gdesc[2u * g + 1u];
            if (gm != 0u) {
                fprog = a.pool + gbase;
                rprog = reinterpret_cast<const RevIns *>(a.pool + gbase + program_stride(gm));
                n = gm;
            }
        }
    }

    template <bool PUSH>
    TSI_DI uint32_t forward(const KArgs &a, Iv x, Iv y, Iv z, Iv t, float2 *stk, Iv &field, Slots &Rlo, Slots &Rhi) {
        Rlo[a.arg_slot[0]] = x.lo; Rhi[a.arg_slot[0]] = x.hi;
        Rlo[a.arg_slot[1]] = y.lo; Rhi[a.arg_slot[1]] = y.hi;
        Rlo[a.arg_slot[2]] = z.lo; Rhi[a.arg_slot[2]] = z.hi;
        Rlo[a.arg_slot[3]] = t.lo; Rhi[a.arg_slot[3]] = t.hi;
        ch.word = 0u;
        ch.n = 0u;
        const uint32_t c = choices ? run_forward<PUSH, true>(fprog, n, ps, Rlo, Rhi, stk, ch)
                                   : run_forward<PUSH, false>(fprog, n, ps, Rlo, Rhi, stk, ch);
        field = I(Rlo[a.out_slot], Rhi[a.out_slot]);
        return c;
    }

    TSI_DI uint32_t reverse(const KArgs &a, const float2 *stk, Iv sens[4], Slots &Rlo, Slots &Rhi) {
        const uint32_t ops = run_reverse(rprog, n, ps, Rlo, Rhi, stk);
        for (int k = 0; k < 4; ++k) sens[k] = I(Rlo[k], Rhi[k]);
        return ops;
    }

    TSI_DI uint32_t records() const { return n; }
    TSI_DI bool choices_recorded() const { return choices; }

    TSI_DI void liveness(const KArgs &a, uint32_t *rec) { live_walk(fprog, n, a.out_slot, ch, a.nchunks, rec, a.prune_cap); }
};

}  // namespace tsi
