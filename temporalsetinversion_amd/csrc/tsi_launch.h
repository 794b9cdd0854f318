// tsi_launch.h -- host-side launch sequences of the frame kernels (tsi_kernels.hip), for the
// runtime.  Kept apart from tsi_kernels.h, whose text is part of every generated tape kernel's
// source: a change here does not invalidate compiled tape kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "tsi_kernels.h"

namespace tsi {

struct FrameLaunch {
    int levels;
    float time;
    uint32_t grid_fwd, grid_rev, grid_misc;
    void *tape_eval;   // the tape's compiled level kernel (hipFunction_t), or null: k_eval
};

// e_eval0 / e_eval1 (may be null) bracket the level loop; e_level (may be null) holds an event
// pair per level, recorded right before and after that level's evaluation launches.
int launch_frame(const KArgs &a, const FrameLaunch &fl, hipStream_t s, hipEvent_t e_eval0, hipEvent_t e_eval1,
                 const hipEvent_t *e_level);
int launch_reset(const KArgs &a, const float box[6], uint64_t nodes, hipStream_t s);
int launch_enqueue_root(const KArgs &a, hipStream_t s);
int launch_trace(const KArgs &a, const TraceArgs &t, hipStream_t s);
int launch_snapshot(const KArgs &a, float4 *s_st, float4 *s_blo, float4 *s_bhi, float4 *s_grad, uint64_t nodes,
                    uint32_t grid, hipStream_t s);
int launch_rgba8(const float4 *in, uchar4 *out, uint32_t n, hipStream_t s);

}  // namespace tsi
