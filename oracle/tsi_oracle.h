/*
 * tsi_oracle.h -- CPU restatement of the temporal set-inversion path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (temporalsetinversion_amd/,
 * include/) may include, link or call this code.  It is loaded by tests/,
 * by __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline
 * leg.  Every function restates a piece of the reference and cites it
 * (paths relative to the reference checkout):
 *
 *   src/ir.rs            tape builder (.ll -> 32-byte GPUInstruction[])
 *   src/compute.wgsl     interval library 2481-2975, `fix` 3462, `main` 3513-4056,
 *                        `trace` 4189-4403, `get_color_*` 4444-4648, `rt_main` 4650-4688,
 *                        `buf_main` 4690-4707
 *   src/main.rs          first-frame seeding 730-766, frame protocol 639-690, 724-787
 *
 * Arithmetic conventions (the reference leaves them to the Vulkan driver; the
 * oracle and the HIP kernels both follow these):
 *   - IEEE-754 binary32, round-to-nearest-even, no FMA contraction, denormals kept;
 *   - WGSL min/max = IEEE-754-2019 minimumNumber/maximumNumber (a NaN operand
 *     yields the other operand, -0 orders below +0);
 *   - sqrt and division correctly rounded;
 *   - sin/exp/log/atan2/pow are the deterministic tsi_det_* routines of
 *     oracle/detmath.c (written from their series definitions; accuracy is
 *     pinned against numpy float64 in tests/test_oracle_math.py).
 */
#ifndef TSI_ORACLE_H
#define TSI_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 32-byte tape record, identical to ir.rs:88-97 GPUInstruction. */
typedef struct {
    uint32_t index_;
    uint32_t opcode_uses;  /* (op | side<<14) << 16 | uses */
    uint32_t larg_rarg;    /* larg << 16 | rarg */
    float literal;
    float val[2];
    float sens[2];
} orc_instr;

typedef struct { float lo, hi; } orc_iv;

/* Voxel record, identical to main.rs:37-52 / compute.wgsl:3003-3016 (88 B). */
typedef struct {
    float x[2], y[2], z[2];
    float t;
    uint32_t child_count;
    uint32_t children[8];
    float f, dx, dy, dz, time_err, fast;
} orc_voxel;

/* The reference's compile-time PARAMETERS (compute.wgsl:3494-3511) as runtime values. */
typedef struct {
    int32_t localsubdiv;
    float delta;
    int32_t point_sample;
    int32_t swept_volume;
    int32_t lipschitz_corrected;
    int32_t fastenabled;
    float fast_multiplier;
    float swept_window;   /* 1/60 in compute.wgsl:3605 */
    int32_t shading;      /* ORC_SHADE_* (compute.wgsl:3503-3504 normal_viz, fast_normals) */
} orc_params;

/* normal_viz (compute.wgsl:3503): hits show normalize(normal) * 0.5 + 0.5, misses (10, 10, 10, 0),
 * no octree darkening.  With ORC_SHADE_GRADIENT (fast_normals = false, 3504) the normal is the
 * voxel's reverse-pass gradient and every cell takes the reverse pass (3780); otherwise it is
 * finite_diff_normal's value for a tape scene, (1, 1, 1) (4143-4184 `default`). */
enum { ORC_SHADE_NORMALS = 1, ORC_SHADE_GRADIENT = 2, ORC_SHADE_CINEMATIC = 4 };
/* cinematic (compute.wgsl:3499): orc_trace renders get_color_supersample's 8-sample path tracer
 * (4620-4648) over get_color_single's two bounces with sun shadow rays and the sunsky
 * background (4440-4523, 520-640); every cell takes the reverse pass with ORC_SHADE_GRADIENT. */

typedef struct {
    uint64_t cell_evals;     /* lanes that passed the evaluation guard (3629) */
    uint64_t rev_evals;      /* of those, lanes that ran the reverse pass (3780) */
    uint64_t fwd_ops;        /* algorithmic f32 ops, forward */
    uint64_t rev_ops;        /* algorithmic f32 ops, reverse (non-short-circuited) */
    uint64_t parents;        /* non-dealloc queue entries processed */
    uint64_t deallocs;       /* dealloc queue entries processed */
    uint64_t enqueued;       /* children enqueued for subdivision */
    uint32_t allocated_voxels;
    uint32_t status;         /* 0 ok, else ORC_E_* */
} orc_stats;

enum { ORC_OK = 0, ORC_E_PARSE = -1, ORC_E_TAPE_TOO_LONG = -2, ORC_E_POOL = -3, ORC_E_QUEUE = -4, ORC_E_ARG = -5 };

/* ---- tape builder (ir.rs:190-366) ---- */
/* Parses `text` (a clang .ll module holding @_Z1fffff) into out[0..n).  time is
 * the ARGW value written into instruction 3 (ir.rs:348).  Returns ORC_OK, or an
 * error; *n_total receives the untruncated instruction count. */
int orc_parse_ll(const char *text, size_t len, float time, orc_instr *out, uint32_t cap, uint32_t *n_total);

/* ---- deterministic math (detmath.c) ---- */
float orc_det_sinf(float x);
float orc_det_expf(float x);
float orc_det_logf(float x);
float orc_det_atan2f(float y, float x);
float orc_det_powi(float x, int32_t n);
float orc_det_pow045(float x);
float orc_det_cosf(float x);
float orc_det_powf(float x, float y);
float orc_det_acosf(float x);
/* cinematic shading pieces (frame.c): frand/frand0to1 (compute.wgsl:2454-2471) advance *seed;
 * getSampleBiased (4425), sunsky (520) and LinearToSRGB (4600) per channel */
float orc_frand(uint32_t *seed);
float orc_frand0to1(uint32_t *seed);
void orc_sample_biased(const float d[3], float power, uint32_t *seed, float out[3]);
void orc_sunsky(const float view[3], float spx, float spy, float out[3]);
float orc_linear_to_srgb(float x);
float orc_wmin(float a, float b);
float orc_wmax(float a, float b);

/* ---- one cell: forward, reverse, time error (compute.wgsl:3622-3996) ---- */
/* Forward interval op exactly as the `switch (op_top)` of 3642-3766 followed by fix(). */
orc_iv orc_forward_op(uint32_t op, uint32_t side, orc_iv l, orc_iv r, float literal);
/* Forward + optional reverse over a tape.  vals/sens: n entries each (may be NULL).
 * Returns the field interval (value of instruction n-1). */
orc_iv orc_eval_tape(const orc_instr *tape, uint32_t n, const orc_iv args[4], int take_derivative,
                     orc_iv *vals, orc_iv *sens, uint64_t *fwd_ops, uint64_t *rev_ops);
/* Per-opcode algorithmic op counts used as roofline numerators. */
uint32_t orc_fwd_cost(uint32_t op, uint32_t side);
uint32_t orc_rev_cost(uint32_t op, uint32_t side);

/* ---- pruning (choice-reachability from RET) ---- */
/* live[j] = 1 when instruction j is reachable from the output through the
 * MIN/MAX choice edges of `vals`; choice[j] = 1 (left only), 2 (right only),
 * 3 (both) for live MIN/MAX, 0 otherwise. */
void orc_prune(const orc_instr *tape, uint32_t n, const orc_iv *vals, uint8_t *live, uint8_t *choice);

/* ---- whole frames: classification, subdivision, traversal ---- */
typedef struct orc_state orc_state;
orc_state *orc_create(const orc_params *p, const orc_instr *tape, uint32_t n, uint32_t voxel_capacity);
void orc_destroy(orc_state *s);
/* main.rs:730-766: clear voxels/queues, root voxel 0 = box, queue = {0}. */
int orc_reset(orc_state *s, const float root_box[6]);
/* compute.wgsl:4690-4707 buf_main: enqueue(0). */
int orc_enqueue_root(orc_state *s);
/* One dispatch of `main` (compute.wgsl:3513-4056) at uniforms.time = time. */
int orc_set_inversion(orc_state *s, float time, orc_stats *st);
/* rt_main (4650-4688) over a width x height frame.  view: column-major 4x4
 * inverse view matrix (uniforms.rs:9); zdist = res.y / tan(radians(fov)/2)
 * precomputed by the host.  rgba: W*H*4 floats, depth: W*H floats (either may be NULL).
 * row0/rows restrict the work to a band of rows (rows = 0 means all). */
int orc_trace(const orc_state *s, const float view[16], uint32_t width, uint32_t height, float time,
              float zdist, float *rgba, float *depth, uint32_t row0, uint32_t rows, uint64_t *node_visits);
/* Canonical export: every node reachable from the root through child_count != 0,
 * as records {depth, i, j, k, f, t, time_err, fast, has_children, dx, dy, dz}. */
typedef struct {
    uint32_t depth, i, j, k;
    float f, t, time_err, fast;
    uint32_t has_children;
    float dx, dy, dz;
} orc_node;
uint64_t orc_export(const orc_state *s, orc_node *out, uint64_t cap);
const orc_voxel *orc_voxels(const orc_state *s, uint32_t *n_used);

/* ---- float64 ray march of a leaf set (march.c): the independent depth check ---- */
/* nodes: a canonical export (orc_export / tsi_read_nodes), borrowed while the set lives;
 * boxes are root_box cut into 2^depth cells per axis, children widened by pad (1e-4 in the
 * reference's trace, compute.wgsl:4310-4311).  depth_out[r] = distance along the unit ray r
 * to the first entered leaf with f <= 0 reached through children with f <= 0, or -1. */
typedef struct orc_leafset orc_leafset;
orc_leafset *orc_leafset_build(const orc_node *nodes, uint64_t n, const double root_box[6], double pad);
void orc_leafset_free(orc_leafset *s);
void orc_leafset_march(const orc_leafset *s, const double *origins, const double *dirs, uint64_t m,
                       double *depth_out);

#ifdef __cplusplus
}
#endif
#endif
