// tsi_launch.h -- host-side launch sequences of the frame kernels (tsi_kernels.hip), for the
// runtime.  Kept apart from tsi_kernels.h, whose text is part of every generated tape kernel's
// source: a change here does not invalidate compiled tape kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "tsi.h"
#include "tsi_kernels.h"

namespace tsi {

// levels evaluated speculatively at frame start unless $TSI_SPEC_LEVELS says otherwise
constexpr long kSpecLevelsDefault = 7;

// every TSI_SHADE_* bit; tsi_params.shading with any other bit is rejected
constexpr uint32_t kShadeBits = TSI_SHADE_NORMALS | TSI_SHADE_GRADIENT | TSI_SHADE_CINEMATIC;

// The frame's speculative pass (KArgs::spec_*) on a stream of its own: it waits for `wait` (the
// previous frame's levels up to the deepest node it reads, or null), runs k_spec_gather and the
// level kernel at level -1 with `args` (its own saved-operand stack) on `blocks` blocks between
// `begin` and `done`; level 0 of the frame waits for `done`.  The frame records `ready` after
// its level `ready_level`, for the next frame's pass.
struct SpecLaunch {
    hipStream_t stream;
    hipEvent_t wait, begin, done, ready;
    KArgs args;
    uint32_t blocks;
    int ready_level;
};

struct FrameLaunch {
    int levels;
    float time;
    uint32_t grid_fwd, grid_rev, grid_misc;
    void *tape_eval;   // the tape's compiled level kernel (hipFunction_t), or null: k_eval
    const SpecLaunch *spec;   // or null
};

// e_eval0 / e_eval1 (may be null) bracket the level loop; e_level (may be null) holds an event
// pair per level, recorded right before and after that level's evaluation launches.
int launch_frame(const KArgs &a, const FrameLaunch &fl, hipStream_t s, hipEvent_t e_eval0, hipEvent_t e_eval1,
                 const hipEvent_t *e_level);
int launch_reset(const KArgs &a, const float box[6], uint64_t nodes, hipStream_t s);
int launch_enqueue_root(const KArgs &a, hipStream_t s);
// blocks > 0 (with next_quad, a counter zeroed on `s` before): the persistent form on that many
// blocks (k_trace_quads); else one 16x16-pixel block per tile (k_trace)
int launch_trace(const KArgs &a, const TraceArgs &t, hipStream_t s, uint32_t blocks = 0, uint32_t *next_quad = nullptr);
int launch_snapshot(const KArgs &a, float4 *s_st, float4 *s_blo, float4 *s_bhi, float4 *s_grad, uint64_t nodes,
                    uint32_t grid, hipStream_t s);
int launch_rgba8(const float4 *in, uchar4 *out, uint32_t n, hipStream_t s);

// One frame of a batched trace (k_trace_batch): the session's nodes and the frame's arguments.
struct TraceJob {
    const float4 *st, *blo, *bhi, *grad;
    TraceArgs t;
    uint32_t blocks_x;  // ceil(width / 16): 16x16-pixel blocks per row
    uint32_t block0;    // the frame's first block in the launch
};
// (cinematic: every job's session shades with TSI_SHADE_CINEMATIC; no job mixes the two)
int launch_trace_batch(const TraceJob *jobs, uint32_t njobs, uint32_t total_blocks, bool cinematic, hipStream_t s);
// The canonical node records of an octree (tsi_read_nodes): every node reachable from the root
// through nodes with children, keyed by (depth, i, j, k).  Synchronous device reads.
int read_octree(const float4 *st_dev, const float4 *grad_dev, const Counters *cnt_dev, uint64_t nodes, tsi_node *out,
                uint64_t cap, uint64_t *count);
// p[0 .. n) = v (a kernel, not a memset node, so that it also replays correctly in a captured graph)
int launch_fill_u32(uint32_t *p, uint32_t v, uint32_t n, hipStream_t s);

}  // namespace tsi
