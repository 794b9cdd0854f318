// tsi_kernels.h -- device data layout and launcher interface of the set-inversion kernels.
//
// Node pool (structure of arrays in HBM, 64 B per node, 16-byte aligned vector accesses):
//   st[n]   = {f, t, time_err, meta}   meta bit31 = fast, bits 0-30 = child block + 1 (0: leaf)
//   blo[n]  = {x.lo, y.lo, z.lo, 0}    box
//   bhi[n]  = {x.hi, y.hi, z.hi, 0}
//   grad[n] = {dx, dy, dz, 0}          sensitivity midpoints (compute.wgsl:3967-3969)
// Children of a node are one block of 8 consecutive nodes (child index x + 2y + 4z, the
// numbering of compute.wgsl:3576 and 4295).  Node 0 is the root; block 0 holds only it.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#include "tsi_tape.h"

namespace tsi {

constexpr int kMaxLevels = 24;
enum : uint32_t { kErrQueue = 1u, kErrPool = 2u };
enum { kStatCells = 0, kStatRevCells, kStatParents, kStatEnqueued, kStatRevOps, kStatDeallocs, kStatLive,
       kStatFwdOps, kStatPruned, kStatSpec, kNumStats = 10 };

// Liveness record of a frontier entry (input of its children's pruned program), stored
// word-major: word w of entry e at rec[w * prune_cap + e], so that a wavefront's consecutive
// entries write consecutive addresses.  Words: 0 = t-window lo, 1 = t-window hi (of the
// evaluation that produced the record), then nchunks live words, nchunks left-kept words and
// nchunks right-kept words, one bit per full-program record.
constexpr uint32_t kRecHead = 2;

struct Counters {
    uint32_t list_count[kMaxLevels + 2][2];   // frontier sizes per level: [0] forward-only, [1] reverse
    uint32_t dlist_count[kMaxLevels + 2];     // dealloc cascade list sizes
    uint32_t alloc_tickets;                   // blocks handed out this frame
    uint32_t free_top;                        // free-stack entries at frame start
    uint32_t bump;                            // blocks ever taken from the bump region
    uint32_t pending;                         // blocks freed this frame
    uint32_t frame_error;                     // kErr* raised this frame
    uint32_t error;                           // sticky kErr* since reset
    uint32_t root_pending;                    // buf_main has enqueued the root
    uint32_t pad;
    uint32_t pool_used[kMaxLevels + 2];       // program-pool records taken per level
    uint32_t pad2[2];
    uint64_t stats[kNumStats];
};

// A speculatively evaluated cell (speculative evaluation of the shallow levels, below): the
// tape's results for one child of a candidate parent and the inputs they were computed from.
// 64 B, one per lane: a wavefront writes 4 KB contiguous.
struct SpecCell {
    float field_lo, field_hi;
    float sens[8];              // sensitivities of x, y, z, t: lo, hi pairs
    float targ_lo, targ_hi;     // the t-window evaluated
    uint32_t fc, ops;           // algorithmic forward / reverse ops
    uint32_t flags;             // kSpec* below
    uint32_t parent;            // the candidate parent's node
};
enum : uint32_t { kSpecEval = 1u, kSpecRev = 2u, kSpecPfast = 4u, kSpecNew = 8u };

struct KArgs {
    float4 *st, *blo, *bhi, *grad;
    uint32_t *lists[2][2];      // [level parity][forward-only / reverse]
    uint32_t *dlists[2];
    uint32_t *pending_free;
    uint32_t *free_stack;
    Counters *cnt;
    float2 *stack;              // reverse-pass operand stack, stack_len * 64 float2 per wavefront
    const FwdIns *fprog;
    const RevIns *rprog;
    uint32_t nfwd, nrev, stack_len, out_slot;
    uint32_t arg_slot[4];
    uint32_t nblocks, list_cap, dlist_cap;
    float epsilon, delta, fast_multiplier, swept_window;
    int32_t swept, point_sample, lipschitz, fastenabled;
    // normal_viz with fast_normals off (compute.wgsl:3780): every cell takes the reverse pass
    int32_t force_deriv;
    // pruned programs (per-group dense programs specialised on the parents' liveness)
    int32_t prune;
    uint32_t *prune_rec[2][2];  // [level parity][forward-only / reverse], see kRecHead
    uint32_t prune_cap;         // entries per list that get a record
    uint32_t nchunks;           // ceil(nfwd / 32)
    uint32_t *gdesc[2];         // per list: {pool offset, records} for each group of 8 entries
    uint32_t prune_live_pct;    // compiled engine: pruned code at most this % of regions live
    uint32_t *choice;           // per wavefront: choice_len words x 64 lanes
    uint32_t choice_len;
    FwdIns *pool;               // program pool: forward records then reverse records per group
    uint32_t pool_cap;          // records
    // speculative evaluation of the shallow levels: at frame start the nodes of the octree at
    // depths < spec_levels are listed (k_spec_gather) and one launch (level -1) evaluates the
    // children of each from the octree as the frame starts (a leaf's as new cells); a level
    // whose parent carries this frame's tag (bhi[P].w) takes a child's results from its record
    // (blo[P].w) when the inputs match, instead of evaluating the cell again
    uint32_t spec_levels;       // 0: off
    uint32_t spec_cap;          // listed parents at most
    uint32_t spec_tag;          // this frame's tag (nonzero, new every frame)
    uint32_t *spec_list;        // the listed parents (breadth first from the root)
    uint32_t *spec_count;       // how many
    SpecCell *spec_res;         // 8 per listed parent
};

struct TraceArgs {
    float view[16];
    uint32_t width, height;
    float time, zdist;
    float4 *rgba;
    float *depth;
    unsigned long long *visits;
    uint32_t shading;           // TSI_SHADE_* (normal_viz, gradient normals, cinematic)
    const float4 *grad;         // the nodes' dx, dy, dz (gradient normals)
};


}  // namespace tsi
