// tsi_eval.h -- one level of the set inversion for one frontier list (compute.wgsl:3513-4056
// restated level-synchronously): allocation of the children of new parents, child boxes and
// t-windows, tape evaluation through an Engine, time error, fast flag, classification and
// compaction of the next frontier.  A 64-lane wavefront takes 8 parents x 8 child octants
// (lane = 8 * parent slot + octant, octant = x + 2y + 4z as compute.wgsl:3576).
//
// Engine (evaluation of the tape for the wavefront's 64 cells):
//   Engine(a, wave, lane)
//   select(a, which, g, pruned_ok, prec_in, rec_idx, need_choices)  program for group g
//                                   (pruned_ok: every parent has a liveness record and the
//                                   group's cells lie inside their parents; the lane's parent
//                                   record is prec_in + rec_idx; need_choices: some cell may
//                                   write a liveness record its children can use)
//   forward<PUSH>(a, x, y, z, t, stk, field, Rlo, Rhi) -> algorithmic f32 ops; PUSH saves the
//                                   operands the reverse rules need on the wavefront's stack
//                                   (stk: uniform base, row r of lane l at stk[64 r + l])
//   reverse(a, stk, sens[4], Rlo, Rhi) -> algorithmic f32 ops; sensitivities of x, y, z, t
//   (Rlo/Rhi: a register file the engine may use between the two calls)
//   records()                      records run per forward pass
//   choices_recorded()             the forward pass recorded the MIN/MAX choices
//   liveness(a, rec)               the cell's liveness record (kPrune engines)
#pragma once
#include "tsi_wave.h"

namespace tsi {
typedef float Slots[kSlots];
}

namespace tsi {


// One step of the dealloc cascade (compute.wgsl:3533-3543) over list `step` (buffer step & 1):
// every node of a freed block that has children passes its child block on to list step + 1,
// and the block goes to the pending free list.  A wavefront takes 8 blocks (lane = 8 * block
// slot + node).  Returns this lane's count of freed blocks.
TSI_DI uint64_t dealloc_step(const KArgs &a, int step, uint32_t wave, uint32_t nwaves) {
    Counters *cnt = a.cnt;
    const uint32_t count = min(cnt->dlist_count[step], a.dlist_cap);
    const uint32_t *in = a.dlists[step & 1];
    uint32_t *out = a.dlists[(step + 1) & 1];
    const uint32_t lane = lane_id();
    const uint32_t o = lane & 7u;
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
    return freed;
}

// Per-lane tallies of the cells a wavefront processed, summed over the wavefront and added to
// the counters once (flush_tally).
struct LevelTally {
    uint64_t cells = 0, rev = 0, par = 0, enq = 0, rev_ops = 0, fwd_ops = 0, pruned = 0, freed = 0, spec = 0;
};

// Frontier entries of one list at a level (a list that overflowed, flagged by wave_append,
// holds list_cap valid entries).  KA: KArgs, or KArgs in the constant address space (a table
// the kernel only reads; its fields are then scalar loads).
template <class KA>
TSI_DI uint32_t list_entries(const KA &a, int level, int which) {
    return min(a.cnt->list_count[level][which], a.list_cap);
}

// Group g (entries 8g .. 8g + 7) of one frontier list at one level: 8 parents x 8 child octants
// on the wavefront's 64 lanes.  count = list_entries(a, level, REV); stk = the wavefront's
// saved-operand stack (uniform base; lane l uses column l of each row).
// spec (reverse list only): the speculative pass of the frame's start: the entries are the
// listed parents (a.spec_list), every child is evaluated as the octree stands (the children of
// a leaf as new cells), and only its SpecCell record and the parent's tag are written.
template <bool REV, class Engine, class KA>
TSI_DI void eval_group(const KA &a, int level, float time, uint32_t g, uint32_t count, float2 *stk, Engine &E,
                       Slots &Rlo, Slots &Rhi, LevelTally &tl, bool spec = false) {
    const int which = REV ? 1 : 0;
    const bool sp = REV && spec;
    Counters *cnt = a.cnt;
    const uint32_t *list = sp ? a.spec_list : a.lists[level & 1][which];
    uint32_t *next_f = a.lists[(level + 1) & 1][0];
    uint32_t *next_r = a.lists[(level + 1) & 1][1];
    const uint32_t lane = lane_id();
    const uint32_t ps = lane >> 3, o = lane & 7u;
    const uint32_t lx = o & 1u, ly = (o >> 1) & 1u, lz = o >> 2;
    const float eps = a.epsilon;
    const bool prune = Engine::kPrune && a.prune != 0;
    // this level's parents' liveness records (written by the previous level) and the next
    // level's (written here)
    const uint32_t *prec_in = a.prune_rec[level & 1][which];
    uint32_t *prec_f = a.prune_rec[(level + 1) & 1][0];
    uint32_t *prec_r = a.prune_rec[(level + 1) & 1][1];
    {
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
            const uint64_t req = sp ? 0u : ballot(isnew && o == 0);
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
        if (!sp && isnew && ok && blk >= a.nblocks) {
            ok = false;
            if (o == 0) atomicOr(&cnt->frame_error, kErrPool);
        }
        if (!sp && isnew && ok && o == 0) a.st[P].w = __uint_as_float((pmeta & kFastBit) | (blk + 1u));
        if (!sp) tl.par += (active && o == 0) ? 1 : 0;

        const uint32_t C = ok && !(sp && isnew) ? blk * 8u + o : 0u;
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
        // (a node keeps its box while it is allocated: only new nodes write it; the speculative
        // pass writes no node)
        if (!sp && ok && isnew) {
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
        const bool td = ct.lo != ct.hi || a.force_deriv != 0;   // take_derivative (3780)

        // ---- program: the group's pruned program when every evaluated child lies inside its
        //      parent (box and t-window of the evaluation that produced the liveness record),
        //      else the full tape ----
        // every parent of the group has a record (capacities are multiples of 8)
        bool pruned_ok = false;
        const uint32_t rec_idx = active ? pidx : g * 8u;
        if (prune && level > 0 && g * 8u + 8u <= a.prune_cap) {
            const float wlo = __uint_as_float(prec_in[rec_idx]);
            const float whi = __uint_as_float(prec_in[a.prune_cap + rec_idx]);
            const bool inside = targ.lo >= wlo && targ.hi <= whi && cx.lo >= plo.x && cx.hi <= phi.x &&
                                cy.lo >= plo.y && cy.hi <= phi.y && cz.lo >= plo.z && cz.hi <= phi.z;
            pruned_ok = ballot(guard && !inside) == 0;
        }
        // A liveness record of this cell serves its children only if their t-windows will lie
        // inside this cell's: children that do not exist yet, or whose windows will be
        // degenerate, start at `time`; existing children start at their own last evaluation,
        // which must not precede this cell's (it does when this cell was evaluated without
        // being subdivided since).  A cell later found fast also gives degenerate windows, but
        // that is not known yet.
        bool kids_ok = false;
        if (prune) {
            kids_ok = a.swept || (cmeta & kChildMask) == 0 || width / 2.0f <= eps;
            if (guard && !kids_ok) {
                const uint32_t cb = ((cmeta & kChildMask) - 1u) * 8u;
                float tmin = a.st[cb].y;
                for (uint32_t k = 1; k < 8; ++k) tmin = fminf(tmin, a.st[cb + k].y);
                kids_ok = tmin >= targ.lo;
            }
        }
        const bool need_choices = prune && ballot(guard && kids_ok && leaf_w > eps) != 0;
        E.select(a, which, g, pruned_ok, prec_in, rec_idx, need_choices);

        // ---- this cell's speculative results: the frame-start pass evaluated it (its parent
        //      carries this frame's tag) from the same inputs ----
        bool have = false;
        const SpecCell *sr = nullptr;
        if (!sp && level < int(a.spec_levels) && ok && __float_as_uint(phi.w) == a.spec_tag) {
            const uint32_t j = __float_as_uint(plo.w);
            if (j < a.spec_cap) {
                sr = a.spec_res + size_t(j) * 8u + o;
                const uint32_t fl = sr->flags;
                // the tape's results depend only on the box (the parent's: the same node), the
                // t-window and whether the reverse pass ran
                have = guard && sr->parent == P && (fl & kSpecEval) != 0 &&
                       __float_as_uint(sr->targ_lo) == __float_as_uint(targ.lo) &&
                       __float_as_uint(sr->targ_hi) == __float_as_uint(targ.hi) && (!REV || !td || (fl & kSpecRev) != 0);
            }
        }

        // ---- tape evaluation (wave-uniform) ----
        Iv field = I_FROM(0.0f);
        uint32_t fcv = 0;
        const uint64_t need_eval = ballot(guard && !have);
        if (need_eval) {
            fcv = E.template forward<REV>(a, cx, cy, cz, targ, stk, field, Rlo, Rhi);
            if (guard && !have && !sp) {
                tl.fwd_ops += fcv;
                tl.pruned += a.nfwd - E.records();
            }
        }
        if (have) {
            field = I(sr->field_lo, sr->field_hi);
            tl.fwd_ops += sr->fc;
        }
        Iv df = I_FROM(0.0f), df_dt = I_FROM(0.0f);
        float4 gradv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float terr = cst.z;
        if (REV) {
            const uint64_t need_rev = ballot(guard && td && !have);
            Iv sens[4] = {I_FROM(0.0f), I_FROM(0.0f), I_FROM(0.0f), I_FROM(0.0f)};
            uint32_t ops = 0;
            if (need_rev) ops = E.reverse(a, stk, sens, Rlo, Rhi);
            if (have && td) {
                for (int k = 0; k < 4; ++k) sens[k] = I(sr->sens[2 * k], sr->sens[2 * k + 1]);
                ops = sr->ops;
            }
            if (sp) {
                // the record of the speculative pass, and the parent's tag for this frame
                if (ok) {
                    SpecCell r;
                    r.field_lo = field.lo;
                    r.field_hi = field.hi;
                    for (int k = 0; k < 4; ++k) {
                        r.sens[2 * k] = sens[k].lo;
                        r.sens[2 * k + 1] = sens[k].hi;
                    }
                    r.targ_lo = targ.lo;
                    r.targ_hi = targ.hi;
                    r.fc = fcv;
                    r.ops = ops;
                    r.flags = (guard ? kSpecEval : 0u) | (guard && td ? kSpecRev : 0u) | (pfast ? kSpecPfast : 0u) |
                              (isnew ? kSpecNew : 0u);
                    r.parent = P;
                    a.spec_res[size_t(g * 8u + ps) * 8u + o] = r;
                    if (o == 0) {
                        reinterpret_cast<float *>(&a.blo[P])[3] = __uint_as_float(g * 8u + ps);
                        reinterpret_cast<float *>(&a.bhi[P])[3] = __uint_as_float(a.spec_tag);
                    }
                }
                return;
            }
            {
                if (guard && td) {
                    tl.rev_ops += ops;
                    const Iv s0 = sens[0], s1 = sens[1], s2 = sens[2];
                    // the sign of a zero sensitivity depends on which pruned instructions ran;
                    // dx/dy/dz are stored with +0 for any zero
                    gradv = make_float4(MID(s0) + 0.0f, MID(s1) + 0.0f, MID(s2) + 0.0f, 0.0f);   // 3967-3969
                    df_dt = sens[3];                                             // 3972
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
        const bool next_fwd_only = !a.swept && !a.force_deriv &&
                                   ((cmeta & kChildMask) == 0 || (cmeta & kFastBit) || width / 2.0f <= eps);
        const uint32_t at_f = wave_append(enq && next_fwd_only, C, next_f, &cnt->list_count[level + 1][0],
                                          a.list_cap, &cnt->frame_error, kErrQueue);
        const uint32_t at_r = wave_append(enq && !next_fwd_only, C, next_r, &cnt->list_count[level + 1][1],
                                          a.list_cap, &cnt->frame_error, kErrQueue);
        // dropped blocks are freed by the next level's forward-only kernel (the dealloc step
        // below), their descendants by the levels after it
        wave_append(drop_children, old_child - 1u, a.dlists[(level + 1) & 1], &cnt->dlist_count[level + 1], a.dlist_cap,
                    &cnt->frame_error, kErrQueue);

        // ---- liveness of the enqueued cells, for their children's pruned programs ----
        // (a record the children cannot use gets an empty window: no t-window lies inside it)
        if (prune && ballot(enq) != 0) {
            uint32_t *rec = nullptr;
            if (at_f < a.prune_cap) rec = prec_f + at_f;
            if (at_r < a.prune_cap) rec = prec_r + at_r;
            const bool valid = kids_ok && E.choices_recorded();
            if (rec) {
                rec[0] = valid ? __float_as_uint(targ.lo) : 0x7f800000u;
                rec[a.prune_cap] = valid ? __float_as_uint(targ.hi) : 0xff800000u;
            }
            if (!valid) rec = nullptr;
            if (ballot(rec != nullptr) != 0) E.liveness(a, rec);
        }
        tl.cells += guard ? 1 : 0;
        tl.spec += have ? 1 : 0;
        tl.rev += (guard && td) ? 1 : 0;
        tl.enq += enq ? 1 : 0;
    }
}

// Adds a wavefront's tallies to the counters: one atomic per statistic.
TSI_DI void flush_tally(Counters *cnt, const LevelTally &t) {
    const uint64_t n_cells = wave_sum64(t.cells), n_rev = wave_sum64(t.rev), n_par = wave_sum64(t.par),
                   n_enq = wave_sum64(t.enq), rev_ops = wave_sum64(t.rev_ops), fwd_ops = wave_sum64(t.fwd_ops),
                   n_pruned = wave_sum64(t.pruned), n_freed = wave_sum64(t.freed), n_spec = wave_sum64(t.spec);
    if (lane_id() == 0) {
        if (n_cells) atomicAdd((unsigned long long *)&cnt->stats[kStatCells], (unsigned long long)n_cells);
        if (n_rev) atomicAdd((unsigned long long *)&cnt->stats[kStatRevCells], (unsigned long long)n_rev);
        if (n_par) atomicAdd((unsigned long long *)&cnt->stats[kStatParents], (unsigned long long)n_par);
        if (n_enq) atomicAdd((unsigned long long *)&cnt->stats[kStatEnqueued], (unsigned long long)n_enq);
        if (rev_ops) atomicAdd((unsigned long long *)&cnt->stats[kStatRevOps], (unsigned long long)rev_ops);
        if (fwd_ops) atomicAdd((unsigned long long *)&cnt->stats[kStatFwdOps], (unsigned long long)fwd_ops);
        if (n_pruned) atomicAdd((unsigned long long *)&cnt->stats[kStatPruned], (unsigned long long)n_pruned);
        if (n_freed) atomicAdd((unsigned long long *)&cnt->stats[kStatDeallocs], (unsigned long long)n_freed);
        if (n_spec) atomicAdd((unsigned long long *)&cnt->stats[kStatSpec], (unsigned long long)n_spec);
    }
}

// One level of one frontier list: wavefront `wave` of `nwaves` strides over the list; `bwave`
// indexes the wavefront's own buffers (saved-operand stack, choice rows), unique in the launch.
// level -1 (reverse list): the speculative pass over the candidate parents (eval_group).
template <bool REV, class Engine>
TSI_DI void eval_level(const KArgs &a, int level, float time, uint32_t wave_, uint32_t nwaves, uint32_t bwave_) {
    const bool sp = REV && level < 0;
    const uint32_t count = sp ? min(*a.spec_count, a.spec_cap) : list_entries(a, level, REV ? 1 : 0);
    // wave indices made visibly uniform: the grid-stride loop and everything inside it then
    // compile as scalar control flow
    const uint32_t wave = __builtin_amdgcn_readfirstlane(wave_);
    const uint32_t bwave = __builtin_amdgcn_readfirstlane(bwave_);
    float2 *stk = REV ? a.stack + size_t(bwave) * a.stack_len * 64 : nullptr;
    Engine E(a, bwave, lane_id());
    // register file of engines that keep values in a uniformly indexed array (a kernel-local
    // array, so that it is promoted to VGPRs; unused by straight-line engines)
    Slots Rlo, Rhi;
    LevelTally tl;
    for (uint32_t g = wave; g * 8u < count; g += nwaves)
        eval_group<REV, Engine>(a, level, time, g, count, stk, E, Rlo, Rhi, tl, sp);
    // ---- the dealloc cascade's step for the blocks dropped by the previous level and their
    //      descendants (forward-only kernel; the reverse kernel of the level does not touch the
    //      list, which the previous level's kernels finished) ----
    if (!REV && level > 0) tl.freed = dealloc_step(a, level, wave, nwaves);
    flush_tally(a.cnt, tl);
}

// The speculative pass's parents (KArgs::spec_*): every node at depth < spec_levels as the frame
// starts, breadth first from the root (the 8 children of each node with children), by the nt
// threads of one block; the list stops at spec_cap entries.
template <class KA>
TSI_DI void spec_gather(const KA &a, uint32_t t, uint32_t nt) {
    __shared__ uint32_t n_sh, begin_sh, end_sh;
    if (t == 0) {
        a.spec_list[0] = 0u;
        n_sh = 1u;
        begin_sh = 0u;
        end_sh = 1u;
    }
    __syncthreads();
    constexpr uint32_t K = 4;   // entries per thread per step, their loads issued together
    for (uint32_t d = 1; d < a.spec_levels; ++d) {   // the nodes at depth d
        const uint32_t b0 = begin_sh, e0 = end_sh;
        for (uint32_t i0 = b0 + t; i0 < e0; i0 += K * nt) {
            uint32_t P[K], m[K];
            for (uint32_t k = 0; k < K; ++k) {
                const uint32_t i = i0 + k * nt;
                P[k] = i < e0 ? a.spec_list[i] : 0u;
            }
            for (uint32_t k = 0; k < K; ++k) m[k] = i0 + k * nt < e0 ? meta_of(a.st[P[k]]) & kChildMask : 0u;
            for (uint32_t k = 0; k < K; ++k) {
                if (m[k]) {
                    const uint32_t at = atomicAdd(&n_sh, 8u);
                    for (uint32_t o = 0; o < 8u; ++o)
                        if (at + o < a.spec_cap) a.spec_list[at + o] = (m[k] - 1u) * 8u + o;
                }
            }
        }
        __syncthreads();
        if (t == 0) {
            begin_sh = e0;
            end_sh = min(n_sh, a.spec_cap);
        }
        __syncthreads();
    }
    if (t == 0) *a.spec_count = min(n_sh, a.spec_cap);
}

// ---- the frame's first and last steps (thread t of nt threads) ----

// Start of a frame: clear the per-frame counters and, when buf_main has enqueued it
// (root_pending), put the root on the level-0 list its state selects.
TSI_DI void frame_begin(const KArgs &a, uint32_t t, uint32_t nt) {
    Counters *c = a.cnt;
    for (uint32_t i = t; i < uint32_t(kMaxLevels + 2); i += nt) {
        c->list_count[i][0] = 0;
        c->list_count[i][1] = 0;
        c->dlist_count[i] = 0;
    }
    for (uint32_t i = t; i < uint32_t(kNumStats); i += nt) c->stats[i] = 0;
    for (uint32_t i = t; i < uint32_t(kMaxLevels + 2); i += nt) c->pool_used[i] = 0;
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
            const bool fwd_only = !a.swept && !a.force_deriv &&
                                  ((meta & kChildMask) == 0 || (meta & kFastBit) || w <= a.epsilon);
            const int which = fwd_only ? 0 : 1;
            a.lists[0][which][0] = 0u;
            c->list_count[0][which] = 1;
        }
    }
}

// Free-list merge: blocks handed out from the top of the stack are gone, freed blocks go on top.
TSI_DI void frame_end(const KArgs &a, uint32_t t, uint32_t nt) {
    Counters *cnt = a.cnt;
    const uint32_t T = cnt->alloc_tickets, F = cnt->free_top, P = min(cnt->pending, a.nblocks);
    const uint32_t used = T < F ? T : F;
    const uint32_t base = F - used;
    for (uint32_t i = t; i < P && base + i < a.nblocks; i += nt) a.free_stack[base + i] = a.pending_free[i];
}

// After frame_end has finished everywhere: the allocator's new state (one thread).
TSI_DI void frame_commit(const KArgs &a) {
    Counters *cnt = a.cnt;
    const uint32_t T = cnt->alloc_tickets, F = cnt->free_top, P = min(cnt->pending, a.nblocks);
    const uint32_t used = T < F ? T : F;
    uint32_t bump = cnt->bump + (T > F ? T - F : 0u);
    cnt->bump = bump < a.nblocks - 1u ? bump : a.nblocks - 1u;
    cnt->free_top = min(F - used + P, a.nblocks);
    cnt->stats[kStatLive] = uint64_t(cnt->bump) - uint64_t(cnt->free_top);
    cnt->error |= cnt->frame_error;
}

// A level kernel that is one launch for one list: wavefront index and count from the grid.
template <bool REV, class Engine>
TSI_DI void eval_level(const KArgs &a, int level, float time) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    eval_level<REV, Engine>(a, level, time, wave, (gridDim.x * blockDim.x) >> 6, wave);
}

// Both lists of a level in one launch of `blocks` blocks of 4 wavefronts: the number of blocks
// that take the forward-only list (and the level's dealloc step), the rest taking the reverse
// list.  Each list gets the blocks its work fills when both fit, else a share of the grid in
// proportion to its work; every block computes the same split from the counts the previous
// launches wrote.
TSI_DI uint32_t level_split(const KArgs &a, int level, uint32_t blocks) {
    const Counters *cnt = a.cnt;
    uint32_t cf = min(cnt->list_count[level][0], a.list_cap);
    if (level > 0) cf = max(cf, min(cnt->dlist_count[level], a.dlist_cap));
    const uint32_t cr = min(cnt->list_count[level][1], a.list_cap);
    const uint32_t bf = (cf + 31u) / 32u, br = (cr + 31u) / 32u;   // 8 entries per wavefront
    if (bf + br <= blocks) return bf;
    if (bf == 0u) return 0u;
    if (br == 0u) return blocks;
    const uint32_t nf = uint32_t(uint64_t(blocks) * bf / (uint64_t(bf) + br));
    return nf < 1u ? 1u : (nf > blocks - 1u ? blocks - 1u : nf);
}

}  // namespace tsi
