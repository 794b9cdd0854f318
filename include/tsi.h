/*
 * tsi.h -- C ABI of the MI355X temporal set-inversion library (libtsi.so).
 *
 * This is the drop-in boundary for the path the reference dispatches through wgpu in
 * src/main.rs.  Each entry point names the reference code it replaces (paths relative to
 * the reference checkout).  Plain pointers and sizes only; no C++ exceptions, no aborts
 * cross this boundary; every call returns TSI_OK (0) or a negative TSI_E_* code.
 * A tsi_ctx is not thread-safe: one context per host thread.  Calls are blocking, as the
 * reference's per-frame timestamp map is (src/main.rs:799-806).
 */
#ifndef TSI_H
#define TSI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TSI_OK = 0,
    TSI_E_OOM = -1,              /* device or host allocation failed */
    TSI_E_QUEUE_FULL = -2,       /* a frontier list overflowed (the reference silently drops, compute.wgsl:3256) */
    TSI_E_POOL_EXHAUSTED = -3,   /* voxel pool full (the reference does not check, compute.wgsl:3440) */
    TSI_E_TAPE_TOO_LONG = -4,    /* > 4096 instructions (the reference truncates, src/ir.rs:358) */
    TSI_E_HIP = -5,              /* a HIP runtime call failed */
    TSI_E_ARG = -6,              /* invalid argument / malformed tape */
    TSI_E_PARSE = -7,            /* .ll text not understood (src/ir.rs would panic) */
    TSI_E_NO_TAPE = -8,          /* no tape uploaded yet */
    TSI_E_UNSUPPORTED = -9,      /* tape needs more live values than the register file holds */
    TSI_E_BUDGET = -10,          /* served form: request exceeds the HBM admission budget */
    TSI_E_JIT = -11,             /* the tape's evaluation kernels did not compile (see tsi_tape_info) */
    TSI_E_PENDING = -12,         /* frames in flight: collect a tsi_set_inversion_result first */
    TSI_E_SCRATCH = -13          /* larger than a static buffer of the library (see its entry point) */
};

/* src/ir.rs:88-97 GPUInstruction, 32 bytes: opcode_uses = (op | side<<14) << 16 | uses,
 * larg_rarg = larg << 16 | rarg, side = 0 reg/reg, 1 reg/lit, 2 lit/reg, 3 lit/lit. */
typedef struct {
    uint32_t index_;
    uint32_t opcode_uses;
    uint32_t larg_rarg;
    float literal;
    float val[2];
    float sens[2];
} tsi_gpu_instruction;

/* src/uniforms.rs:8-17, 96 bytes.  view = inverse view matrix, column-major (camera.rs:33). */
typedef struct {
    float view[16];
    uint32_t resolution[2];
    float time;
    float lasttime;
    float rendermode;
    float pad[3];
} tsi_uniforms;

/* compute.wgsl:3494-3511 PARAMETERS as runtime values, plus pool sizing. */
typedef struct {
    int32_t localsubdiv;          /* octree depth L: (2^L)^3 leaves over the root box */
    float delta;                  /* temporal error bound */
    int32_t point_sample;         /* evaluate leaf cells at their centre */
    int32_t swept_volume;
    int32_t lipschitz_corrected;
    int32_t fastenabled;
    float fast_multiplier;
    float swept_window;           /* 1/60 in compute.wgsl:3605 */
    uint64_t voxel_capacity;      /* nodes in the pool (0 = default) */
    uint64_t queue_capacity;      /* entries per frontier list (0 = voxel_capacity / 8) */
    uint32_t flags;               /* TSI_FLAG_* */
    uint32_t prune_capacity;      /* frontier entries per list and level that get a pruned
                                     program (0 = min(queue_capacity, 2^19)) */
    uint32_t prune_live_max;      /* compiled engine: a wavefront runs its pruned code when at
                                     most this percentage of the tape's regions is live for it
                                     (0 = 75; 100 = whenever its parents' records apply) */
    uint32_t shading;             /* TSI_SHADE_* (0: the reference's default colouring; other
                                     bits than TSI_SHADE_* give TSI_E_ARG) */
} tsi_params;

/* compute.wgsl:3499, 3503-3504 normal_viz / fast_normals / cinematic as runtime values.
 * TSI_SHADE_NORMALS: hits show normalize(normal) * 0.5 + 0.5, misses (10, 10, 10, 0) before the
 * 0.45 gamma, no octree darkening (4265, 4333-4377, 4510-4514).  With TSI_SHADE_GRADIENT as well
 * (fast_normals = false) the normal is the voxel's reverse-pass gradient (dx, dy, dz) and every
 * cell takes the reverse pass (3780); without it, it is finite_diff_normal's value for a tape
 * scene, (1, 1, 1).  TSI_SHADE_CINEMATIC (cinematic = true): the path tracer of
 * get_color_supersample / get_color_single (4440-4523, 4620-4648): 8 jittered samples per pixel,
 * two bounces drawn from a cosine-power lobe around the hit normal (the normal as above; every
 * cell takes the reverse pass when TSI_SHADE_GRADIENT is set too), a sun shadow ray per bounce,
 * the sunsky background (520-640) for rays that leave after the first bounce, the mean through
 * LinearToSRGB and the 0.45 gamma; alpha is the fraction of samples whose first ray hit.  The
 * lighting is the scene-independent one of 4446-4452.  The depth output is the first sample's
 * first ray.  Tracing costs up to 32 rays per pixel. */
enum { TSI_SHADE_NORMALS = 1u, TSI_SHADE_GRADIENT = 2u, TSI_SHADE_CINEMATIC = 4u };

enum {
    /* Evaluate the full tape in every cell (the reference's behaviour).  Without it, a cell
     * runs the tape pruned on its parent's MIN/MAX choices (compute.wgsl:2884-2903): results
     * are identical for tapes of +, -, *, /literal, sqrt, min, max, abs, neg, floor; for tapes
     * that add sin, cos, exp or log they are identical unless an interval bound is within a
     * few ulp of a MIN/MAX decision; tapes with any other opcode are never pruned. */
    TSI_FLAG_NO_PRUNE = 1u,
    TSI_FLAG_PRUNE_EXACT_ONLY = 2u, /* prune only tapes of the first (exact) kind */
    /* Evaluate tapes with the interpreter (programs in constant memory) instead of compiling
     * each uploaded tape into straight-line evaluation kernels (which prune by regions of
     * 16 records, see prune_live_max). */
    TSI_FLAG_INTERPRET = 4u,
    /* Do not evaluate the shallow levels speculatively.  By default the compiled engine, when
     * the tape is not pruned, evaluates at each frame's start, in one launch, the children of
     * every node of the first levels (depth < 7, $TSI_SPEC_LEVELS sets another count, 0 none)
     * as the octree stands (a leaf's children as new cells), where each of those levels
     * otherwise runs one wavefront's dependent chain through the tape on an idle device; a
     * level then takes a cell's results from that launch when they were computed from the
     * cell's inputs of this frame (same box and t-window) and evaluates the others.  Results
     * and statistics are the same either way; the cells of parents that are not enqueued this
     * frame are extra work. */
    TSI_FLAG_NO_SPECULATE = 8u
};

/* Per-frame statistics.  fwd_ops/rev_ops count algorithmic f32 interval operations. */
typedef struct {
    double gpu_ms;                /* device time of the set-inversion dispatches */
    double eval_ms;               /* of which: tape-evaluation kernels */
    uint64_t cell_evals;          /* child cells that passed the evaluation guard (compute.wgsl:3629) */
    uint64_t rev_evals;           /* of those, cells that ran the reverse pass (compute.wgsl:3780) */
    uint64_t fwd_ops;             /* of the records actually run (pruned programs run fewer) */
    uint64_t rev_ops;
    uint64_t parents;             /* non-dealloc work items (parents subdivided) */
    uint64_t deallocs;            /* node blocks freed by the dealloc cascade */
    uint64_t enqueued;            /* cells enqueued for subdivision */
    uint64_t allocated_nodes;     /* live pool nodes after the frame */
    uint64_t eval_launches;       /* tape-evaluation kernel launches */
    uint64_t pruned_instr;        /* sum over evaluated cells of instructions skipped by pruning */
    uint64_t prune_launches;      /* group-program generator launches */
    double eval_kernel_ms;        /* the evaluation launches alone: device time between events
                                     recorded right before and after each level's launches
                                     (eval_ms also holds the gaps between levels) */
    uint64_t spec_cells;          /* of cell_evals, cells whose results came from the frame's
                                     speculative pass (TSI_FLAG_NO_SPECULATE) */
} tsi_stats;

/* Canonical node record (see tsi_read_nodes). */
typedef struct {
    uint32_t depth, i, j, k;
    float f, t, time_err, fast;
    uint32_t has_children;
    float dx, dy, dz;
} tsi_node;

typedef struct tsi_ctx tsi_ctx;

/* Defaults of compute.wgsl:3494-3511 with the given depth. */
void tsi_default_params(tsi_params *p, int32_t localsubdiv);

/* State::new buffers and pipelines (src/main.rs:97-496): creates the device pool, frontier
 * lists and stream on the current HIP device. */
int tsi_create(const tsi_params *params, tsi_ctx **out);
void tsi_destroy(tsi_ctx *ctx);

/* First-frame tape upload (src/main.rs:742-746): validates and lowers n GPUInstruction
 * records (n <= 4096 else TSI_E_TAPE_TOO_LONG).  The host copy is kept for tsi_read_tape. */
int tsi_upload_tape(tsi_ctx *ctx, const tsi_gpu_instruction *tape, uint32_t n);

/* first_eval block (src/main.rs:730-766): clear pool and queues, root node = root_box
 * (x.lo, y.lo, z.lo, x.hi, y.hi, z.hi), enqueue the root. */
int tsi_reset(tsi_ctx *ctx, const float root_box[6]);

/* State::update uniform write (src/main.rs:503). */
int tsi_set_uniforms(tsi_ctx *ctx, const tsi_uniforms *u);

/* One `main` dispatch (src/main.rs:776-814, compute.wgsl:3513-4056) at uniforms.time.
 * Blocking.  *gpu_ms (may be NULL) receives the device time, like State::compute's return. */
int tsi_set_inversion(tsi_ctx *ctx, double *gpu_ms, tsi_stats *stats);

/* The same frame without waiting for it (no counterpart in the reference, whose frame loop
 * waits on every dispatch): the frame is queued behind the previous work of the context and
 * its counters are copied to pinned host memory in stream order.  At most two frames may be in
 * flight (TSI_E_PENDING otherwise); tsi_set_inversion_result waits for the oldest one and
 * returns what tsi_set_inversion would have (its return code is that frame's).  While frames are
 * in flight, tsi_set_inversion, tsi_upload_tape and tsi_reset return TSI_E_PENDING. */
int tsi_set_inversion_async(tsi_ctx *ctx);
int tsi_set_inversion_result(tsi_ctx *ctx, double *gpu_ms, tsi_stats *stats);

/* rt_main (src/main.rs:649-656, compute.wgsl:4650-4688): traces the frame into the
 * library's device pixel buffer and copies it to rgba_out (W*H*4 floats, row 0 on top)
 * and depth_out (W*H floats, -1 = miss); either may be NULL. */
int tsi_trace(tsi_ctx *ctx, float *rgba_out, float *depth_out);

/* The same trace without waiting (no counterpart in the reference, whose render call is
 * synchronous): the nodes in use are copied to a device-side snapshot in stream order after
 * the last set inversion (the count is read on the device: no host wait), and a stream of its
 * own traces the snapshot while the caller goes on with the next frame's set inversion.  At most
 * two traces are in flight, each with its own stream, snapshot (48 bytes per node of
 * voxel_capacity, allocated at the first trace that uses it) and frame, so a trace may start
 * while the last wavefronts of the previous one still run; a third waits on the host for the
 * oldest.  tsi_trace_wait waits for every trace in flight and copies the last frame out (rgba /
 * depth may be NULL). */
int tsi_trace_async(tsi_ctx *ctx);
int tsi_trace_wait(tsi_ctx *ctx, float *rgba_out, float *depth_out);
/* Device time and node visits summed over the traces completed since the context was made. */
int tsi_trace_totals(tsi_ctx *ctx, double *trace_ms, uint64_t *node_visits, uint64_t *traces);

/* Device pointers of the frame written by the last tsi_trace (valid until the second
 * tsi_trace after it or tsi_destroy: the two traces in flight write two frames in turn), for
 * callers that keep the frame on the GPU. */
int tsi_frame_device(tsi_ctx *ctx, const float **rgba_dev, const float **depth_dev);

/* buf_main (src/main.rs:683-690, compute.wgsl:4690-4707): re-enqueue the root so the
 * next tsi_set_inversion re-walks the octree. */
int tsi_enqueue_root(tsi_ctx *ctx);

/* Debug/parity export (cf. the tape/voxel dumps of src/main.rs:822-876): every node
 * reachable from the root through nodes with children, as tsi_node records keyed by
 * (depth, i, j, k).  Writes min(cap, count) records; *count receives the total. */
int tsi_read_nodes(tsi_ctx *ctx, tsi_node *out, uint64_t cap, uint64_t *count);

/* The uploaded tape as 32-byte records (src/main.rs:822-876 tape dump). */
int tsi_read_tape(tsi_ctx *ctx, tsi_gpu_instruction *out, uint32_t cap, uint32_t *n);

/* The frame dump of src/main.rs:693-715: the last traced frame converted to RGBA8 (clamp to
 * [0,1], x255, round half away from zero, as image 0.24 `into_rgba8`; row 0 = top), copied to
 * rgba8_out (W*H*4 bytes), or written as a PNG file.  tsi_write_png writes any RGBA8 image. */
int tsi_frame_rgba8(tsi_ctx *ctx, uint8_t *rgba8_out);
int tsi_save_frame_png(tsi_ctx *ctx, const char *path);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): tsi_frame_rgba8 with
 * its RGBA8 staging buffer in the library's static device memory (TSI_STATIC_RGBA8_MAX_PIXELS
 * pixels, 32 MiB, one copy per GPU, held until the process ends) instead of a hipMalloc per call.
 * Allocates nothing.  Frames of more pixels return TSI_E_SCRATCH.  All calls on one GPU share
 * that buffer: calls that overlap on one GPU (several contexts or host threads) race. */
#define TSI_STATIC_RGBA8_MAX_PIXELS (1u << 23)
int tsi_frame_rgba8_noalloc(tsi_ctx *ctx, uint8_t *rgba8_out);
int tsi_write_png(const char *path, const uint8_t *rgba8, uint32_t width, uint32_t height);

/* Device time of the last tsi_trace dispatch and the octree nodes it visited (child-state
 * loads of the traversal; no counterpart in the reference, which has no timer on rt_main). */
int tsi_last_trace_info(tsi_ctx *ctx, double *trace_ms, uint64_t *node_visits);

/* How the uploaded tape is evaluated (no counterpart in the reference, which interprets). */
typedef struct {
    uint32_t n;                   /* tape records */
    uint32_t stack;               /* saved operand intervals per lane (reverse pass) */
    uint32_t slots;               /* interpreter register-file slots used */
    uint32_t prune_class;         /* 0 never pruned, 1 pruned up to ulp-level decisions, 2 exact */
    uint32_t engine;              /* 0 interpreter, 1 compiled straight-line kernels */
    uint32_t cache_hit;           /* compiled kernels came from the code-object cache */
    uint32_t specialized;         /* the kernels assume every cell box inside [-5,5]^3 and every
                                     t-window inside [-2, 64] (rebuilt without it otherwise) */
    uint32_t fix_kept;            /* records whose fix is evaluated (the rest provably never clamp) */
    double compile_ms;            /* hiprtc compile (or cache lookup) wall time */
    char log[1024];               /* compiler log of a failed compile (truncated) */
} tsi_tape_info;
int tsi_tape_info_get(tsi_ctx *ctx, tsi_tape_info *out);

/* Compiles a tape's evaluation kernels into the code-object cache without a device or
 * context (no counterpart in the reference): a later tsi_upload_tape of the same tape with the
 * same pruning flags (TSI_FLAG_NO_PRUNE, TSI_FLAG_PRUNE_EXACT_ONLY) then only loads them.  Cache directory: $TSI_JIT_CACHE, else $XDG_CACHE_HOME/tsi-jit, else
 * ~/.cache/tsi-jit. */
int tsi_precompile_tape(const tsi_gpu_instruction *tape, uint32_t n, uint32_t flags, tsi_tape_info *out);

/* The source of the evaluation kernels tsi_precompile_tape / tsi_upload_tape would compile for a
 * tape under these flags, without compiling it (no counterpart in the reference, which interprets
 * the tape): writes min(cap - 1, length) characters and a NUL to out (may be NULL); *len receives
 * the length; *info (may be NULL) the lowering's figures (n, stack, slots, prune_class). */
int tsi_tape_kernel_source(const tsi_gpu_instruction *tape, uint32_t n, uint32_t flags, char *out, size_t cap,
                           size_t *len, tsi_tape_info *info);

/* src/ir.rs IR::new + parseLLVMIR + to_gpu: builds the tape from .ll text holding
 * @_Z1fffff.  time is written into instruction 3 (ARGW).  *n receives the full
 * instruction count (also when > cap or > 4096). */
int tsi_tape_from_ll(const char *text, size_t len, float time, tsi_gpu_instruction *out, uint32_t cap,
                     uint32_t *n);

/* z = res.y / tan(radians(90)/2) of compute.wgsl:4592 for a frame of the given height,
 * as the library computes it. */
float tsi_camera_zdist(uint32_t height);

/* ---- batched and served forms (no counterpart in the reference, which renders one scene) ----
 *
 * A batch holds several tapes and any number of sessions.  A session is what a tsi_ctx holds for
 * one scene -- an octree over a root box, evaluated with one of the batch's tapes and its own
 * parameters -- kept in the batch's device arena, so that its octree stays resident between
 * frames (the temporal reuse of compute.wgsl:3603-3614 within delta).  tsi_batch_submit runs one
 * frame of each listed session: one persistent launch walks every session's octree levels
 * together, its wavefronts taking work units (8 frontier entries of one session) from a device
 * queue, then one launch traces every frame that asks for one (16x16-pixel blocks of all the
 * frames).  Each session's frame is the one a tsi_ctx with the same
 * tape and parameters computes: the same octree nodes, statistics and pixels, bit for bit.
 * Sessions are admitted by their device footprint against the batch's arena (TSI_E_BUDGET).
 * A batch is not thread-safe; submits are serialised (one in flight). */
typedef struct tsi_batch tsi_batch;

/* The batch's kernels: tapes[i] (lens[i] records) is tape i of tsi_session_desc.tape.  The
 * module is compiled (or found in the code-object cache) at creation; arena_bytes of device
 * memory are reserved for sessions (0: 1/2 of the device's free memory). */
typedef struct {
    uint64_t arena_bytes;         /* admission budget: sessions' pools, lists and frames */
    uint32_t max_sessions;        /* sessions open at once (0 = 1024, the most) */
    uint32_t waves_per_simd;      /* persistent grid: 1..4 wavefronts per SIMD (0 = 2) */
    uint32_t claim_units;         /* work units (8 frontier entries each) per queue claim (0 = 1) */
} tsi_batch_config;

typedef struct {
    uint32_t tape;                /* index into the batch's tapes */
    tsi_params params;            /* localsubdiv, delta, point_sample, swept_volume, swept_window,
                                     lipschitz_corrected, fastenabled, fast_multiplier, shading,
                                     voxel_capacity (0 = 2^22 nodes), queue_capacity; flags,
                                     prune_* are ignored: batches evaluate the full tape */
    float root_box[6];            /* x.lo, y.lo, z.lo, x.hi, y.hi, z.hi (inside [-5,5]^3) */
    uint32_t max_width, max_height;   /* largest frame it traces (0, 0: 1920 x 1080) */
} tsi_session_desc;

/* One frame of one session: uniforms.time = time; traced (width x height, view) unless width
 * or height is 0.  flags: TSI_FRAME_RESET = start from an empty octree (tsi_reset inside the
 * launch: a request that needs no earlier frame). */
enum { TSI_FRAME_RESET = 1u };
typedef struct {
    uint32_t session;
    float time;
    uint32_t width, height;
    float view[16];
    uint32_t flags;
} tsi_frame_req;

typedef struct {
    int32_t status;               /* TSI_OK, TSI_E_QUEUE_FULL or TSI_E_POOL_EXHAUSTED */
    uint32_t pad;
    tsi_stats stats;              /* as tsi_set_inversion's (gpu_ms, eval_ms: the whole batch's) */
    uint64_t node_visits;         /* of its trace */
} tsi_frame_result;

typedef struct {
    uint32_t tapes, grid_blocks, waves_per_simd, stack_rows;
    uint32_t sessions_open, cache_hit;
    uint32_t max_waves_per_simd;  /* most the kernel's registers and the stack allow */
    uint32_t claim_units;
    double compile_ms;
    uint64_t arena_bytes, arena_used;
    uint64_t stack_bytes;
    char log[1024];
} tsi_batch_info;

int tsi_batch_create(const tsi_gpu_instruction *const *tapes, const uint32_t *lens, uint32_t ntapes,
                     const tsi_batch_config *cfg, tsi_batch **out);
void tsi_batch_destroy(tsi_batch *b);
/* Compiles a batch module into the code-object cache without a device (tsi_precompile_tape's
 * counterpart): a later tsi_batch_create with the same tapes only loads it. */
int tsi_batch_precompile(const tsi_gpu_instruction *const *tapes, const uint32_t *lens, uint32_t ntapes,
                         tsi_batch_info *out);
int tsi_batch_info_get(tsi_batch *b, tsi_batch_info *out);
/* Launch shape between submits (the autotuner's knobs): the non-zero fields of cfg among
 * waves_per_simd (<= tsi_batch_info.max_waves_per_simd) and claim_units replace the batch's;
 * arena_bytes and max_sessions are ignored. */
int tsi_batch_tune(tsi_batch *b, const tsi_batch_config *cfg);
/* Bytes of arena a session of this description takes (its admission footprint). */
int tsi_batch_footprint(tsi_batch *b, const tsi_session_desc *d, uint64_t *bytes);
/* Admits a session (TSI_E_BUDGET when its footprint does not fit in the arena's free space) and
 * resets its octree to the root box (tsi_reset). */
int tsi_batch_open(tsi_batch *b, const tsi_session_desc *d, uint32_t *session);
int tsi_batch_close(tsi_batch *b, uint32_t session);
/* tsi_reset of one session. */
int tsi_batch_reset(tsi_batch *b, uint32_t session);
/* One frame of each listed session (each at most once): set inversion, trace, then the root's
 * re-enqueue (buf_main), queued on the batch's stream without waiting.  TSI_E_PENDING while the
 * previous submit has not been collected. */
int tsi_batch_submit(tsi_batch *b, const tsi_frame_req *reqs, uint32_t n);
/* TSI_OK when the submitted frames are done (or none is in flight), TSI_E_PENDING while they run;
 * does not wait. */
int tsi_batch_query(tsi_batch *b);
/* Waits for the submitted frames; results[i] (may be NULL) for reqs[i]; set_ms / trace_ms (may
 * be NULL): device time of the set-inversion launch and of the trace launch. */
int tsi_batch_wait(tsi_batch *b, tsi_frame_result *results, double *set_ms, double *trace_ms);
/* tsi_read_nodes / tsi_trace_wait's frame of one session (its last traced frame). */
int tsi_batch_read_nodes(tsi_batch *b, uint32_t session, tsi_node *out, uint64_t cap, uint64_t *count);
int tsi_batch_frame(tsi_batch *b, uint32_t session, float *rgba_out, float *depth_out);

/* Parameters of a context between frames (no counterpart in the reference, whose PARAMETERS are
 * shader constants): localsubdiv, delta, point_sample, swept_volume, swept_window,
 * lipschitz_corrected, fastenabled and fast_multiplier take the new values; voxel_capacity,
 * queue_capacity, flags, prune_* and shading must equal the context's (TSI_E_ARG otherwise).
 * With frames in flight: TSI_E_PENDING. */
int tsi_set_params(tsi_ctx *ctx, const tsi_params *params);

/* Text of a TSI_E_* code. */
const char *tsi_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif
