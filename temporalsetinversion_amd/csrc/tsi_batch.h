// tsi_batch.h -- the persistent set-inversion kernel of a batch: one launch runs one frame of
// every session in the batch (sessions of different tapes, octree depths, t-windows and
// parameters), level by level, driven by a device work queue.
//
// The reference's `main` is one persistent dispatch fed by a broker queue of parent voxels
// (compute.wgsl:3513-3556, enqueue/dequeue 3249-3282), for one scene.  Here the grid is
// resident for the whole run (a cooperative launch: every block on a CU at once) and walks the
// octree levels of all sessions together:
//
//   phase "begin"      every session's frame_begin and the node list of its speculative pass
//                      (spec_gather; block-strided over sessions)
//   phase -1           the speculative pass: groups of 8 listed nodes of every session
//                      (eval_group's speculative mode), when the sessions have one
//   phase d = 0 .. D+1 work units of level d, claimed by wavefronts from a counter until none
//                      is left: for each session, groups of 8 entries of its reverse list, of its
//                      forward-only list (levels below its depth), and of its dealloc list
//                      (step d of the cascade, d >= 1; steps D and D+1 are the cascade's tail)
//   phase "end"        every session's frame_end, then frame_commit
//
// with a grid barrier between phases.  A unit is exactly what the per-level kernels do with one
// group (eval_group / dealloc_step, tsi_eval.h), on the session's own pool and lists, so a
// session's frame is the same whether it runs alone or in a batch.  Within a phase, units of
// every session are in one queue ordered reverse units first (the costly ones), then
// forward-only, then dealloc units: a wavefront that finishes takes the next unit of any
// session, so a level's work fills the machine as long as the batch has any.
//
// Tape: each session names a tape of the module (Tapes::group<REV> switches on it); the tape
// code is the per-tape straight-line code of tsi_jit.cpp, one namespace per tape.
#pragma once
#include "tsi_eval.h"

namespace tsi {

constexpr uint32_t kBatchMaxSessions = 1024;
enum : uint32_t { kErrBarrier = 4u };   // a grid barrier timed out (the run is abandoned)

struct BatchSession {
    KArgs a;           // the session's pool, lists and counters (a.stack unused: the run's stack)
    float time;        // the frame's time (uniforms.time)
    uint32_t tape;     // the session's tape in the module
    int32_t levels;    // localsubdiv
    uint32_t reset;    // start from an empty octree over `box` (tsi_reset) before the frame
    float box[6];      // root box (x.lo, y.lo, z.lo, x.hi, y.hi, z.hi)
    uint32_t pad[2];
};

// Device control block of a run, cleared (by a fill kernel) before the launch.  The barrier's
// counters sit in lines of their own (64 bytes apart).
struct BatchCtl {
    uint32_t bar_count[16];               // grid barrier: arrivals since the run began
    uint32_t bar_gen[16];                 // barriers completed
    uint32_t error[16];                   // kErrBarrier
    uint32_t claim[kMaxLevels + 4];       // work-queue heads, one per level phase
};

struct BatchArgs {
    const BatchSession *sess;
    BatchCtl *ctl;
    Counters *out;          // each session's counters after the frame (read back by the host)
    float2 *stack;          // saved-operand stacks, stack_len rows x 64 lanes per wavefront
    uint32_t stack_len;
    uint32_t nsess;
    int32_t phases;         // levels + 2 (the deepest session's depth, then the cascade's tail)
    int32_t first_phase;    // -1: the sessions evaluate their shallow levels speculatively, else 0
    uint32_t claim_units;   // units a wavefront takes per claim
    uint64_t wait_ticks;    // longest wait at a grid barrier, in s_memrealtime ticks (100 MHz)
};

// Grid-wide barrier over the resident grid (all blocks run at once: cooperative launch): the
// first thread of each block arrives on one counter with a release (its block's writes, which
// __syncthreads has completed, are written back from the XCD's L2), the last arrival acquires
// them all and publishes the next generation with a release, and the others poll the generation
// and acquire.  (A two-level tree of counters measured slower at one block per CU: the fences,
// not the atomics, set a barrier's cost.)  The wait is bounded: past wait_ticks the run is
// flagged and every block leaves at its next barrier.  Returns false when the run is to be
// abandoned (the same answer in every thread of a block).
TSI_DI bool grid_sync(const BatchArgs &b, uint32_t &gen) {
    __shared__ uint32_t ok_sh;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t target = (gen + 1u) * gridDim.x;
        const uint32_t old = __hip_atomic_fetch_add(&b.ctl->bar_count[0], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1u == target) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(&b.ctl->bar_gen[0], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(&b.ctl->bar_gen[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gen + 1u) {
                __builtin_amdgcn_s_sleep(1);
                if (wall_clock64() - t0 > b.wait_ticks) {
                    atomicOr(&b.ctl->error[0], kErrBarrier);
                    break;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        ok_sh = __hip_atomic_load(&b.ctl->error[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u ? 1u : 0u;
    }
    __syncthreads();
    gen += 1u;
    return ok_sh != 0u;
}

// Work units of session s in phase d: reverse groups, forward-only groups, dealloc groups
// (phase -1: the speculative pass's groups, as reverse units).
TSI_DI void batch_units(const BatchSession &S, int d, uint32_t &nr, uint32_t &nf, uint32_t &nd) {
    nr = nf = nd = 0u;
    if (d < 0) {
        if (S.a.spec_levels) nr = (min(*S.a.spec_count, S.a.spec_cap) + 7u) / 8u;
        return;
    }
    if (d < S.levels) {
        nr = (list_entries(S.a, d, 1) + 7u) / 8u;
        nf = (list_entries(S.a, d, 0) + 7u) / 8u;
    }
    if (d >= 1) nd = (min(S.a.cnt->dlist_count[d], S.a.dlist_cap) + 7u) / 8u;
}

// One dealloc unit: group g of the cascade's list `step` (dealloc_step's loop body).
template <class KA>
TSI_DI uint64_t dealloc_group(const KA &a, int step, uint32_t g) {
    Counters *cnt = a.cnt;
    const uint32_t count = min(cnt->dlist_count[step], a.dlist_cap);
    const uint32_t *in = a.dlists[step & 1];
    uint32_t *out = a.dlists[(step + 1) & 1];
    const uint32_t lane = lane_id();
    const uint32_t o = lane & 7u;
    const uint32_t idx = g * 8u + (lane >> 3);
    const bool active = idx < count;
    const uint32_t blk = active ? in[idx] : 0u;
    const uint32_t m = active ? meta_of(a.st[blk * 8u + o]) & kChildMask : 0u;
    wave_append(m != 0, m - 1u, out, &cnt->dlist_count[step + 1], a.dlist_cap, &cnt->frame_error, kErrQueue);
    wave_append(active && o == 0, blk, a.pending_free, &cnt->pending, a.nblocks, &cnt->frame_error, kErrQueue);
    return (active && o == 0) ? 1u : 0u;
}

// The persistent kernel's body.  Tapes::group<REV>(tape, a, level, time, g, count, stk, tally)
// runs eval_group with the tape's engine.  One block per CU, of 4, 8, 12 or 16 wavefronts (one
// to four per SIMD).
template <class Tapes>
TSI_DI void batch_run(const BatchArgs &b) {
    // prefix sums of the phase's units over sessions, in queue order: [0, n) reverse units of
    // sessions 0.., then forward-only, then dealloc; pre[k] = units before segment k
    __shared__ uint32_t pre[3 * kBatchMaxSessions + 1];
    uint32_t gen = 0;
    const uint32_t tid = threadIdx.x, nb = gridDim.x;
    const uint32_t ns = b.nsess < kBatchMaxSessions ? b.nsess : kBatchMaxSessions;

    // ---- reset (tsi_reset: main.rs:730-766 first-frame seeding) and frame begin ----
    for (uint32_t s = blockIdx.x; s < ns; s += nb) {
        const BatchSession &S = b.sess[s];
        if (S.reset && tid < 8u) {
            S.a.st[tid] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (tid == 0) {
                S.a.blo[0] = make_float4(S.box[0], S.box[1], S.box[2], 0.0f);
                S.a.bhi[0] = make_float4(S.box[3], S.box[4], S.box[5], 0.0f);
                S.a.grad[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                Counters *c = S.a.cnt;
                c->root_pending = 1u;
                c->alloc_tickets = 0u;
                c->free_top = 0u;
                c->bump = 0u;
                c->pending = 0u;
                c->error = 0u;
                c->frame_error = 0u;
            }
        }
        __syncthreads();
        frame_begin(S.a, tid, blockDim.x);
        if (S.a.spec_levels) spec_gather(S.a, tid, blockDim.x);
        __syncthreads();
    }
    if (!grid_sync(b, gen)) return;

    const uint32_t lane = lane_id();
    const uint32_t wpb = blockDim.x >> 6;   // wavefronts per block
    const uint32_t nwaves = nb * wpb;
    const uint32_t bwave = __builtin_amdgcn_readfirstlane(blockIdx.x * wpb + (tid >> 6));
    float2 *stk = b.stack + size_t(bwave) * b.stack_len * 64u;
    const uint32_t cu = b.claim_units;
    // the first round's units go to the blocks in turn (wavefront w of block i takes slot
    // w * nb + i), so that a phase of a few units runs on as many CUs
    const uint32_t slot = __builtin_amdgcn_readfirstlane((tid >> 6) * nb + blockIdx.x);

    for (int d = b.first_phase; d < b.phases; ++d) {
        // ---- the phase's unit table (every block builds the same one) ----
        for (uint32_t s = tid; s < ns; s += blockDim.x) {
            uint32_t nr, nf, nd;
            batch_units(b.sess[s], d, nr, nf, nd);
            pre[1u + s] = nr;
            pre[1u + ns + s] = nf;
            pre[1u + 2u * ns + s] = nd;
        }
        if (tid == 0) pre[0] = 0u;
        __syncthreads();
        // inclusive scan of pre[1 .. 3 ns] (Hillis-Steele over the block, in place)
        const uint32_t m = 3u * ns;
        for (uint32_t off = 1; off < m; off <<= 1) {
            uint32_t v[12];
            uint32_t k = 0;
            for (uint32_t i = tid; i < m && k < 12u; i += blockDim.x, ++k) v[k] = i >= off ? pre[1u + i - off] : 0u;
            __syncthreads();
            k = 0;
            for (uint32_t i = tid; i < m && k < 12u; i += blockDim.x, ++k) pre[1u + i] += v[k];
            __syncthreads();
        }
        const uint32_t total = pre[m];

        // ---- units: in the first round wavefront slot i takes unit i; then wavefronts claim cu
        //      units at a time from the phase's counter, which hands out units from nwaves on
        //      (a phase of no more units than wavefronts takes no atomic) ----
        for (uint32_t base = slot, take = 1u; base < total; take = cu) {
            const uint32_t end = min(base + take, total);
            for (uint32_t u = base; u < end; ++u) {
                // segment j with pre[j] <= u < pre[j + 1]
                uint32_t lo = 0, hi = m;
                while (hi - lo > 1u) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (pre[mid] <= u) lo = mid; else hi = mid;
                }
                const uint32_t j = __builtin_amdgcn_readfirstlane(lo);
                const uint32_t kind = j / ns, s = j - kind * ns;
                const uint32_t g = __builtin_amdgcn_readfirstlane(u - pre[j]);
                // the session table is written before the launch and only read here: a unit's
                // kernel arguments come from scalar loads, as a launch's own arguments do
                const auto &S = as_const(b.sess[s]);
                const auto &a = S.a;
                const uint32_t tape = S.tape;
                const float time = S.time;
                LevelTally tl;
                if (kind == 0u)
                    Tapes::template group<true>(tape, a, d, time, g,
                                                d < 0 ? min(*a.spec_count, a.spec_cap) : list_entries(a, d, 1), stk, tl);
                else if (kind == 1u)
                    Tapes::template group<false>(tape, a, d, time, g, list_entries(a, d, 0), stk, tl);
                else
                    tl.freed = dealloc_group(a, d, g);
                flush_tally(a.cnt, tl);
            }
            if (nwaves >= total) break;
            uint32_t next = 0;
            if (lane == 0) next = atomicAdd(&b.ctl->claim[d + 1], cu);
            base = __builtin_amdgcn_readfirstlane(bcast(next, 0)) + nwaves;
        }
        if (!grid_sync(b, gen)) return;
    }

    // ---- frame end: free-list merge (all threads of a block per session), then commit ----
    for (uint32_t s = blockIdx.x; s < ns; s += nb) frame_end(b.sess[s].a, tid, blockDim.x);
    if (!grid_sync(b, gen)) return;
    // commit, the frame's counters for the host, and buf_main: the root re-enqueued for the
    // session's next frame (compute.wgsl:4690-4707)
    for (uint32_t s = blockIdx.x * blockDim.x + tid; s < ns; s += nb * blockDim.x) {
        const KArgs &a = b.sess[s].a;
        frame_commit(a);
        b.out[s] = *a.cnt;
        a.cnt->root_pending = 1u;
    }
}

}  // namespace tsi
