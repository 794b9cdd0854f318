/*
 * frame.c -- classification/subdivision, time error, octree trace and frame protocol of
 * the oracle.  TEST INFRASTRUCTURE ONLY (see tsi_oracle.h).
 *
 * orc_set_inversion restates one dispatch of compute.wgsl `main` (3513-4056) executed by a
 * single workgroup: the broker queue (3249-3282) becomes a FIFO, mallocV/freeV (3437-3460)
 * keep their free-list FIFO and bump pointer, and the eight lanes of a parent run one after
 * another.  Per-frame results do not depend on that order: a child reads only its parent's
 * box and `fast` flag (both final before the parent is enqueued) and its own previous state.
 * orc_trace restates rt_main -> get_color_supersample -> get_color_single -> trace for the
 * non-cinematic, non-normal_viz configuration (compute.wgsl:3499-3502), scene without
 * per-scene eye tweaks.
 */
#include "tsi_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef orc_iv iv;
#define BIG 1000000.0f
#define FAILURE 0xFFFFFFFFu

struct orc_state {
    orc_params p;
    orc_instr *tape;
    uint32_t n;
    orc_voxel *vox;
    uint32_t cap;
    uint32_t start_of_empty_voxel;      /* Memory.start_of_empty_voxel */
    uint32_t allocated;                 /* Memory.allocated_voxels */
    uint32_t *q;                        /* work queue FIFO */
    uint64_t qh, qt, qcap;
    uint32_t *fq;                       /* voxel free-list FIFO */
    uint64_t fh, ft, fcap;
    int status;
    iv *vals, *sens;
};

static inline float wmin(float a, float b) { return orc_wmin(a, b); }
static inline float wmax(float a, float b) { return orc_wmax(a, b); }
static inline iv I(float lo, float hi) { iv r; r.lo = lo; r.hi = hi; return r; }
static inline iv I_FROM(float a) { return I(a, a); }
static inline int is_nan_w(float a) { return !(a < 0.0f || a > 0.0f || a == 0.0f); }
static inline float DIAMETER(iv a) { return a.hi - a.lo; }
static inline float MID(iv a) { return (a.lo + a.hi) / 2.0f; }
static inline int I_CONTAINS_X(iv a, float x) { return a.lo <= x && x <= a.hi; }
static inline int I_CONTAINS_ZERO(iv a) { return a.lo <= 0.0f && 0.0f <= a.hi; }
static iv MUL_IF(iv a, float b) { return b >= 0.0f ? I(a.lo * b, a.hi * b) : I(a.hi * b, a.lo * b); }
static iv ADD_II(iv a, iv b) { return I(a.lo + b.lo, a.hi + b.hi); }
static iv SQRT_I(iv a) { return I(sqrtf(wmax(a.lo, 0.0f)), sqrtf(wmax(a.hi, 0.0f))); }
static iv ABS_I(iv a)
{
    if (I_CONTAINS_ZERO(a)) return I(0.0f, wmax(-a.lo, a.hi));
    if (a.hi < 0.0f) return I(-a.hi, -a.lo);
    return a;
}
static iv SQR_I(iv a)
{
    float x = a.lo * a.lo, y = a.hi * a.hi;
    if (I_CONTAINS_ZERO(a)) return I(0.0f, wmax(x, y));
    return I(wmin(x, y), wmax(x, y));
}
static iv CAP_II(iv a, iv b)
{
    if (a.hi < b.lo || b.hi < a.lo) return I(INFINITY, -INFINITY);
    return I(wmax(a.lo, b.lo), wmin(a.hi, b.hi));
}
static iv DIV_II(iv a, iv b)
{
    float x = a.lo / b.lo, y = a.lo / b.hi, z = a.hi / b.lo, w = a.hi / b.hi;
    return I(wmin(wmin(x, y), wmin(z, w)), wmax(wmax(x, y), wmax(z, w)));
}
static inline float wclamp(float e, float lo, float hi) { return wmin(wmax(e, lo), hi); }

/* ---------------- queues and pool ---------------- */

static int qpush(uint32_t **buf, uint64_t *h, uint64_t *t, uint64_t *cap, uint32_t v)
{
    if (*t - *h == *cap) {
        uint64_t nc = *cap * 2;
        uint32_t *nb = (uint32_t *)malloc(nc * sizeof(uint32_t));
        if (!nb) return 0;
        for (uint64_t i = *h; i < *t; i++) nb[i - *h] = (*buf)[i % *cap];
        free(*buf);
        *buf = nb;
        *t -= *h;
        *h = 0;
        *cap = nc;
    }
    (*buf)[*t % *cap] = v;
    (*t)++;
    return 1;
}

static int enqueue(orc_state *s, uint32_t v)          /* compute.wgsl:3249 */
{
    if (!qpush(&s->q, &s->qh, &s->qt, &s->qcap, v)) { s->status = ORC_E_QUEUE; return 0; }
    return 1;
}

static uint32_t mallocV(orc_state *s)                 /* compute.wgsl:3437 */
{
    uint32_t a;
    if (s->fh < s->ft) {
        a = s->fq[s->fh % s->fcap];
        s->fh++;
    } else {
        a = ++s->start_of_empty_voxel;
        if (a >= s->cap) { s->status = ORC_E_POOL; a = s->cap - 1; }
    }
    orc_voxel *v = &s->vox[a];
    v->child_count = 0;
    v->f = 1.0f;
    v->t = 0.0f;
    v->time_err = 0.0f;
    v->fast = 0.0f;
    v->dx = v->dy = v->dz = 0.0f;   /* not reset by compute.wgsl:3437-3449; zeroed here so a reused
                                       voxel never shows a previous occupant's values */
    s->allocated++;
    return a;
}

static void freeV(orc_state *s, uint32_t a)           /* compute.wgsl:3452 */
{
    orc_voxel *v = &s->vox[a];
    v->child_count = 0;
    v->f = 0.0f;
    v->t = 0.0f;
    v->time_err = 0.0f;
    v->fast = 0.0f;
    s->allocated--;
    if (!qpush(&s->fq, &s->fh, &s->ft, &s->fcap, a)) s->status = ORC_E_QUEUE;
}

orc_state *orc_create(const orc_params *p, const orc_instr *tape, uint32_t n, uint32_t voxel_capacity)
{
    if (!p || !tape || n < 5 || voxel_capacity < 16) return NULL;
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    if (!s) return NULL;
    s->p = *p;
    s->n = n;
    s->tape = (orc_instr *)malloc(n * sizeof(orc_instr));
    s->cap = voxel_capacity;
    s->vox = (orc_voxel *)calloc(voxel_capacity, sizeof(orc_voxel));
    s->qcap = 1024;
    s->q = (uint32_t *)malloc(s->qcap * sizeof(uint32_t));
    s->fcap = 1024;
    s->fq = (uint32_t *)malloc(s->fcap * sizeof(uint32_t));
    s->vals = (iv *)malloc(n * sizeof(iv));
    s->sens = (iv *)malloc(n * sizeof(iv));
    if (!s->tape || !s->vox || !s->q || !s->fq || !s->vals || !s->sens) { orc_destroy(s); return NULL; }
    memcpy(s->tape, tape, n * sizeof(orc_instr));
    return s;
}

void orc_destroy(orc_state *s)
{
    if (!s) return;
    free(s->tape); free(s->vox); free(s->q); free(s->fq); free(s->vals); free(s->sens);
    free(s);
}

/* main.rs:730-766 */
int orc_reset(orc_state *s, const float root_box[6])
{
    memset(s->vox, 0, (size_t)s->cap * sizeof(orc_voxel));
    s->qh = s->qt = 0;
    s->fh = s->ft = 0;
    s->start_of_empty_voxel = 0;
    s->allocated = 0;
    s->status = ORC_OK;
    orc_voxel *r = &s->vox[0];
    r->x[0] = root_box[0]; r->x[1] = root_box[3];
    r->y[0] = root_box[1]; r->y[1] = root_box[4];
    r->z[0] = root_box[2]; r->z[1] = root_box[5];
    enqueue(s, 0u);
    return ORC_OK;
}

int orc_enqueue_root(orc_state *s)                     /* compute.wgsl:4705-4707 */
{
    return enqueue(s, 0u) ? ORC_OK : ORC_E_QUEUE;
}

/* One workgroup iteration on parent voxel P (compute.wgsl:3558-4053). */
static void process_parent(orc_state *s, uint32_t P, float time, float epsilon, orc_stats *st)
{
    const orc_params *p = &s->p;
    uint32_t ci[8];
    int new_voxel = s->vox[P].child_count == 0u;                          /* 3564 */
    if (new_voxel)
        for (int j = 0; j < 8; j++) ci[j] = mallocV(s);                   /* 3566-3570 */
    for (uint32_t lane = 0; lane < 8; lane++) {
        uint32_t lx = lane & 1u, ly = (lane >> 1) & 1u, lz = lane >> 2;
        uint32_t cur = new_voxel ? ci[lane] : s->vox[P].children[lane];  /* 3574-3577 */
        orc_voxel *pv = &s->vox[P];
        pv->child_count = 1u;                                             /* 3579 */
        pv->children[lane] = cur;                                         /* 3580 */
        iv px = I(pv->x[0], pv->x[1]), py = I(pv->y[0], pv->y[1]), pz = I(pv->z[0], pv->z[1]);
        float dx = DIAMETER(px) / 2.0f, dy = DIAMETER(py) / 2.0f, dz = DIAMETER(pz) / 2.0f;   /* 3582 */
        float width = wmax(dx, wmax(dy, dz));
        iv cx = I(px.lo + dx * (float)lx, px.lo + dx * (float)(1u + lx));   /* 3587 */
        iv cy = I(py.lo + dy * (float)ly, py.lo + dy * (float)(1u + ly));
        iv cz = I(pz.lo + dz * (float)lz, pz.lo + dz * (float)(1u + lz));
        orc_voxel *cv = &s->vox[cur];
        cv->x[0] = cx.lo; cv->x[1] = cx.hi;                               /* 3598 */
        cv->y[0] = cy.lo; cv->y[1] = cy.hi;
        cv->z[0] = cz.lo; cv->z[1] = cz.hi;
        iv ct;
        if (new_voxel || pv->fast == 1.0f || width <= epsilon) {         /* 3603 */
            if (p->swept_volume) ct = I(time - p->swept_window, time);
            else ct = I_FROM(time);
            cv->fast = 0.0f;
        } else {
            ct = I(cv->t, time);
        }
        cv->t = time;                                                     /* 3614 */
        if (p->point_sample && width <= epsilon) {                        /* 3616 */
            cx = I_FROM(MID(cx));
            cy = I_FROM(MID(cy));
            cz = I_FROM(MID(cz));
            width = 0.0f;
        }
        if (!(new_voxel || cv->f >= 0.0f || cv->child_count == 0u))      /* 3629 */
            continue;
        iv args[4] = {cx, cy, cz, I(wmin(ct.lo, ct.hi), wmax(ct.lo, ct.hi))};
        int take_derivative = ct.lo != ct.hi ||                           /* 3780 */
                              ((p->shading & (ORC_SHADE_NORMALS | ORC_SHADE_CINEMATIC)) &&
                               (p->shading & ORC_SHADE_GRADIENT));
        iv field = orc_eval_tape(s->tape, s->n, args, take_derivative, s->vals, s->sens,
                                 st ? &st->fwd_ops : NULL, st ? &st->rev_ops : NULL);
        if (st) { st->cell_evals++; if (take_derivative) st->rev_evals++; }
        iv df = I_FROM(0.0f), df_dt = I_FROM(0.0f);
        if (take_derivative) {
            iv *sens = s->sens;
            /* 3967-3969.  "+ 0.0f" turns a -0 into +0: the sign of a zero sensitivity is not
             * part of the result (it depends on the order in which zero contributions are
             * summed, which evaluating a pruned tape changes). */
            cv->dx = MID(sens[0]) + 0.0f;
            cv->dy = MID(sens[1]) + 0.0f;
            cv->dz = MID(sens[2]) + 0.0f;
            df_dt = sens[3];                                              /* 3972 */
            float dt = DIAMETER(s->vals[3]);
            df = MUL_IF(df_dt, dt);
            iv W = MUL_IF(I(-1.0f, 1.0f), DIAMETER(field));
            df = CAP_II(df, W);
            if (p->lipschitz_corrected && ABS_I(df).hi != 0.0f) {         /* 3980 */
                float lb = ABS_I(SQRT_I(ADD_II(SQR_I(sens[0]), ADD_II(SQR_I(sens[1]), SQR_I(sens[2]))))).lo;
                if (!is_nan_w(lb)) df = DIV_II(df, I_FROM(lb));
            }
            if (df.lo == -INFINITY || df.hi == INFINITY || is_nan_w(df.lo) || is_nan_w(df.hi))   /* 3993 */
                df = I_FROM(p->delta);
            cv->time_err = cv->time_err + wclamp(ABS_I(df).hi, 0.0f, p->delta);                   /* 3996 */
        }
        if (!p->swept_volume && p->fastenabled)                           /* 4003 */
            cv->fast = ABS_I(df).hi >= p->delta * p->fast_multiplier ? 1.0f : 0.0f;
        float thick = 0.0f;
        if (I_CONTAINS_X(field, p->delta * thick)) {                      /* 4011 */
            cv->f = 0.0f;
            if (width > epsilon) {
                if (new_voxel || is_nan_w(df_dt.lo) || is_nan_w(df_dt.hi) || fabsf(df_dt.lo) == INFINITY ||
                    fabsf(df_dt.hi) == INFINITY || pv->fast == 1.0f || ct.lo == ct.hi ||
                    cv->time_err >= p->delta || cv->child_count == 0u) {
                    cv->time_err = 0.0f;
                    enqueue(s, cur);
                    if (st) st->enqueued++;
                }
            }
        } else {
            if (field.hi <= p->delta * thick) cv->f = -1.0f;              /* 4028 */
            else if (!p->swept_volume) cv->f = 1.0f;
            if (!p->swept_volume) {                                       /* 4036 */
                if (cv->child_count != 0u) {
                    cv->child_count = 0u;
                    for (int j = 0; j < 8; j++) enqueue(s, cv->children[j] | (1u << 31));
                }
            }
        }
    }
}

int orc_set_inversion(orc_state *s, float time, orc_stats *st)
{
    float epsilon = 10.0f * (float)ldexp(1.0, -s->p.localsubdiv);        /* 3521 */
    if (st) memset(st, 0, sizeof *st);
    while (s->qh < s->qt && s->status == ORC_OK) {
        uint32_t deq = s->q[s->qh % s->qcap];
        s->qh++;
        uint32_t idx = deq & ~(1u << 31);
        if (deq >> 31) {                                                  /* 3533 */
            if (s->vox[idx].child_count != 0u) {
                s->vox[idx].child_count = 0u;
                for (int j = 0; j < 8; j++) enqueue(s, s->vox[idx].children[j] | (1u << 31));
            }
            freeV(s, idx);
            if (st) st->deallocs++;
            continue;
        }
        if (st) st->parents++;
        process_parent(s, idx, time, epsilon, st);
    }
    if (st) { st->allocated_voxels = s->allocated; st->status = (uint32_t)s->status; }
    return s->status;
}

/* ---------------- trace (rt_main) ---------------- */

typedef struct { float x, y, z; } v3;
static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscale(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vsadd(v3 a, float s) { return V(a.x + s, a.y + s, a.z + s); }
static inline float vlength(v3 a) { return sqrtf((a.x * a.x + a.y * a.y) + a.z * a.z); }
static inline float max_component(v3 a) { return wmax(a.x, wmax(a.y, a.z)); }
static inline float min_component(v3 a) { return wmin(a.x, wmin(a.y, a.z)); }
static inline float wsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
static int32_t f2i_sat(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)x;
}

#define TR_EPS 0.0001f                    /* compute.wgsl:4140 */
#define DARKEN 0.98f                      /* compute.wgsl:3507 */

/* compute.wgsl:4059-4070 */
static v3 efficient_slabs(v3 p0, v3 p1, v3 o, v3 inv)
{
    v3 t0 = vmul(vsub(p0, o), inv);
    v3 t1 = vmul(vsub(p1, o), inv);
    v3 tmin = V(wmin(t0.x, t1.x), wmin(t0.y, t1.y), wmin(t0.z, t1.z));
    v3 tmax = V(wmax(t0.x, t1.x), wmax(t0.y, t1.y), wmax(t0.z, t1.z));
    float d1 = max_component(tmin), d2 = min_component(tmax);
    return V((d1 <= d2 && d2 >= 0.0f) ? 1.0f : 0.0f, wmax(d1, 0.0f), wmax(d2, 0.0f));
}

static inline float mylerp(float a, float b, float t)        /* 4116: mix(a, b, clamp(t, 0, 1)) */
{
    float tt = wmin(wmax(t, 0.0f), 1.0f);
    return a * (1.0f - tt) + b * tt;
}

/* normalize (WGSL), as v / length(v) */
static inline v3 vnormalize(v3 v) { float l = vlength(v); return V(v.x / l, v.y / l, v.z / l); }

/* compute.wgsl:4189-4403 (normal_viz and cinematic from s->p.shading).  *nrm receives
 * normal_at_hit: finite_diff_normal's (1, 1, 1) for a tape scene, or the voxel's gradient with
 * ORC_SHADE_GRADIENT (fast_normals = false, 4333-4342); (0, 0, 0) on a miss. */
static void trace_ray(const orc_state *s, v3 ro, v3 rd, float time, v3 *col, float *dist, uint64_t *visits,
                      v3 *nrm)
{
    const int normal_viz = (s->p.shading & ORC_SHADE_NORMALS) != 0;
    const int cinematic = (s->p.shading & ORC_SHADE_CINEMATIC) != 0;
    v3 normal_at_hit = V(0.0f, 0.0f, 0.0f);
    const orc_voxel *vox = s->vox;
    uint32_t stack[16];
    uint32_t sp = 0;
    v3 c = V(1.0f, 1.0f, 1.0f);
    v3 o2 = ro;
    int intersected = 0;
    v3 inv = V(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
#define PUSH(v) do { if (sp != 16u) stack[sp++] = (v); } while (0)
#define POP() (sp == 0u ? FAILURE : stack[--sp])
#define PEAK() (sp == 0u ? FAILURE : stack[sp - 1])
    PUSH(0u);
    uint32_t region = PEAK();
    v3 p0 = V(vox[region].x[0], vox[region].y[0], vox[region].z[0]);
    v3 p1 = V(vox[region].x[1], vox[region].y[1], vox[region].z[1]);
    v3 hit = efficient_slabs(p0, p1, o2, inv);
    if (nrm) *nrm = normal_at_hit;
    if (hit.x == 0.0f) { *col = c; *dist = -1.0f; return; }
    float sms = 0.0f;
    for (uint32_t iter = 0; iter < 5000u; iter++) {
        if (intersected) break;
        if (sp == 0u) { *col = c; *dist = -1.0f; return; }
        region = POP();
        if (hit.x == 0.0f) {
            uint32_t top = PEAK();
            if (top != FAILURE) {
                p0 = V(vox[top].x[0], vox[top].y[0], vox[top].z[0]);
                p1 = V(vox[top].x[1], vox[top].y[1], vox[top].z[1]);
                hit = efficient_slabs(p0, p1, o2, inv);
            }
            continue;
        }
        v3 ext = vsub(p1, p0);
        v3 rop = V(((o2.x + (hit.y + TR_EPS) * rd.x) - p0.x) * 2.0f / ext.x,
                   ((o2.y + (hit.y + TR_EPS) * rd.y) - p0.y) * 2.0f / ext.y,
                   ((o2.z + (hit.y + TR_EPS) * rd.z) - p0.z) * 2.0f / ext.z);
        v3 rdp = V(rd.x * (2.0f / ext.x), rd.y * (2.0f / ext.y), rd.z * (2.0f / ext.z));
        int32_t mx = f2i_sat(floorf(rop.x)), my = f2i_sat(floorf(rop.y)), mz = f2i_sat(floorf(rop.z));
        v3 dd = V(1.0f / fabsf(rdp.x), 1.0f / fabsf(rdp.y), 1.0f / fabsf(rdp.z));
        v3 sg = V(wsign(rdp.x), wsign(rdp.y), wsign(rdp.z));
        int32_t sx = f2i_sat(sg.x), sy = f2i_sat(sg.y), sz = f2i_sat(sg.z);
        v3 sd = V(((sg.x * ((float)mx - rop.x)) + sg.x * 0.5f + 0.5f) * dd.x,
                  ((sg.y * ((float)my - rop.y)) + sg.y * 0.5f + 0.5f) * dd.y,
                  ((sg.z * ((float)mz - rop.z)) + sg.z * 0.5f + 0.5f) * dd.z);
        float mkx = 0.0f, mky = 0.0f, mkz = 0.0f;
        v3 old = o2;
        for (int i = 0; i < 32; i++) {
            if (!cinematic && !normal_viz && max_component(vsub(p1, p0)) < 1.0f) c = vscale(c, DARKEN);   /* 4265 */
            float t = ((vlength(V(mkx * (sd.x - dd.x), mky * (sd.y - dd.y), mkz * (sd.z - dd.z))) + hit.y) + TR_EPS) + sms;
            o2 = vadd(old, vscale(rd, t));
            sms = 0.0f;
            if (t >= hit.z - TR_EPS) {                                                /* 4279 */
                o2 = vadd(old, vscale(rd, hit.z));
                uint32_t top = PEAK();
                if (top != FAILURE) {
                    p0 = V(vox[top].x[0], vox[top].y[0], vox[top].z[0]);
                    p1 = V(vox[top].x[1], vox[top].y[1], vox[top].z[1]);
                    hit = efficient_slabs(p0, p1, o2, inv);
                }
                break;
            }
            int32_t childindex = (int32_t)((uint32_t)mx + 2u * (uint32_t)my + 4u * (uint32_t)mz);
            if (childindex < 0 || childindex >= 8) break;                             /* 4296 */
            uint32_t child = vox[region].children[childindex];
            if (visits) (*visits)++;
            if (vox[child].f <= 0.0f) {
                v3 tp0 = V(vox[child].x[0] - TR_EPS, vox[child].y[0] - TR_EPS, vox[child].z[0] - TR_EPS);
                v3 tp1 = V(vox[child].x[1] + TR_EPS, vox[child].y[1] + TR_EPS, vox[child].z[1] + TR_EPS);
                v3 th = efficient_slabs(tp0, tp1, o2, inv);
                if (th.x == 1.0f) {
                    if (vox[child].child_count != 0u) {
                        PUSH(region);
                        PUSH(child);
                        p0 = tp0;
                        p1 = tp1;
                        hit = th;
                        break;
                    } else {
                        intersected = 1;
                        normal_at_hit = (s->p.shading & ORC_SHADE_GRADIENT)
                                            ? V(vox[child].dx, vox[child].dy, vox[child].dz)
                                            : V(1.0f, 1.0f, 1.0f);
                        if (normal_viz) {                                     /* 4333-4377 */
                            v3 n = normal_at_hit;
                            v3 nv = vnormalize(V(n.x * 1.0f, n.y * 1.0f, n.z * 1.0f));
                            c = V(nv.x * 0.5f + 0.5f, nv.y * 0.5f + 0.5f, nv.z * 0.5f + 0.5f);
                            break;
                        }
                        float ts = time - vox[child].t;
                        float tt = ts * 60.0f;
                        if (cinematic) {   /* movement (0.7, 1e-4, 1e-4), base 0.1: 4346-4374 */
                            v3 m1 = V(1.0f - 0.7f, 1.0f - 0.0001f, 1.0f - 0.0001f);
                            v3 b1 = V(1.0f - 0.1f, 1.0f - 0.1f, 1.0f - 0.1f);
                            c = V(c.x - mylerp(m1.x, b1.x, tt), c.y - mylerp(m1.y, b1.y, tt), c.z - mylerp(m1.z, b1.z, tt));
                        } else {
                            c = V(c.x - mylerp(0.0f, 0.0f, tt), c.y - mylerp(1.0f, 0.0f, tt), c.z - mylerp(1.0f, 0.0f, tt));
                        }
                        break;
                    }
                }
            }
            int bx = sd.x <= wmin(sd.y, sd.z), by = sd.y <= wmin(sd.z, sd.x), bz = sd.z <= wmin(sd.x, sd.y);
            mkx = bx ? 1.0f : 0.0f; mky = by ? 1.0f : 0.0f; mkz = bz ? 1.0f : 0.0f;
            sd = V(sd.x + mkx * dd.x, sd.y + mky * dd.y, sd.z + mkz * dd.z);
            mx = (int32_t)((uint32_t)mx + (uint32_t)(bx * sx));
            my = (int32_t)((uint32_t)my + (uint32_t)(by * sy));
            mz = (int32_t)((uint32_t)mz + (uint32_t)(bz * sz));
        }
    }
    c = V(wmin(c.x, 1.0f), wmin(c.y, 1.0f), wmin(c.z, 1.0f));
    c = V(wmax(c.x, 0.0f), wmax(c.y, 0.0f), wmax(c.z, 0.0f));
    float d = fabsf(vlength(vsub(o2, ro)));
    if (!intersected) d = -1.0f;
    *col = c;
    *dist = d;
    if (nrm) *nrm = normal_at_hit;
#undef PUSH
#undef POP
#undef PEAK
}

/* ---------------- cinematic shading ---------------- */

static inline float vdot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b)
{
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

#define PI_W 3.14159274101257324f         /* compute.wgsl:2476 PI as f32 */

/* frand0to1 (compute.wgsl:2466-2471): the LCG step seed * 747796405 + 1, then PCG's
 * RXS-M-XS output permutation, mapped to [0, 1] by / 4294967295.0 (2^32 as f32). */
float orc_frand0to1(uint32_t *seed)
{
    *seed = *seed * 747796405u + 1u;
    uint32_t w = ((*seed >> ((*seed >> 28u) + 4u)) ^ *seed) * 277803737u;
    w = (w >> 22u) ^ w;
    return (float)w / 4294967295.0f;
}

/* frand (2454-2459): the same draw mapped to [-1, 1] */
float orc_frand(uint32_t *seed) { return orc_frand0to1(seed) * 2.0f - 1.0f; }

/* ortho (4405-4423): a vector perpendicular to v */
static v3 ortho(v3 v)
{
    return fabsf(v.x) > fabsf(v.z) ? V(-v.y, v.x, 0.0f) : V(0.0f, -v.z, v.y);
}

/* getSampleBiased (4425-4434): cosine-power lobe of exponent `power` around d */
static v3 sample_biased(v3 d, float power, uint32_t *seed)
{
    v3 dir = vnormalize(d);
    v3 o1 = vnormalize(ortho(dir));
    v3 o2 = vnormalize(vcross(dir, o1));
    float rx = orc_frand0to1(seed);
    float ry = orc_frand0to1(seed);
    rx = rx * 2.0f * PI_W;
    ry = orc_det_powf(ry, 1.0f / (power + 1.0f));
    float om = sqrtf(1.0f - ry * ry);
    float cx = orc_det_cosf(rx) * om, sy = orc_det_sinf(rx) * om;
    return vadd(vadd(vscale(o1, cx), vscale(o2, sy)), vscale(dir, ry));
}

void orc_sample_biased(const float d[3], float power, uint32_t *seed, float out[3])
{
    v3 r = sample_biased(V(d[0], d[1], d[2]), power, seed);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* fromSpherical (494-518) */
static v3 from_spherical(float px, float py)
{
    return V(orc_det_cosf(px) * orc_det_sinf(py), orc_det_sinf(px) * orc_det_sinf(py), orc_det_cosf(py));
}

/* RayleighPhase (429-436), hgPhase (458-476) with pow(x, 2.0) as x * x */
static float rayleigh_phase(float c) { return (3.0f / (16.0f * 3.1415927410125732f)) * (1.0f + c * c); }
static float hg_phase(float c, float g)
{
    return (1.0f / (4.0f * 3.1415927410125732f)) *
           ((1.0f - g * g) / orc_det_powf((1.0f - (2.0f * g) * c) + g * g, 1.5f));
}

/* SunIntensity (478-492), with the sky's sunIntensity = 1000 (427) */
static float sun_intensity(float zc)
{
    return 1000.0f * wmax(0.0f, 1.0f - orc_det_expf(-((1.6110731611233942f - orc_det_acosf(zc)) / 1.5f)));
}

/* totalMie (437-456), pow(x, 2.0) as x * x */
static v3 total_mie(v3 lambda, v3 K, float T)
{
    float c = (0.20000000298023224f * T) * 9.99999983775159e-18f;
    float k = (0.4339999854564667f * c) * 3.1415927410125732f;
    const float tp = 2.0f * 3.1415927410125732f;
    v3 q = V(tp / lambda.x, tp / lambda.y, tp / lambda.z);
    return V((k * (q.x * q.x)) * K.x, (k * (q.y * q.y)) * K.y, (k * (q.z * q.z)) * K.z);
}

/* sunsky (520-640; Fragmentarium's Sunsky.frag), turbidity = sunSize = 1 (423-424).  The
 * sun disk it computes is not in its result. */
static v3 sunsky(v3 view, float spx, float spy)
{
    float sadc = orc_det_cosf((1.0f * 3.1415927410125732f) / 180.0f);
    v3 sd = vnormalize(from_spherical((spx - 0.0f) * 6.28000020980835f, (spy - 0.5f) * 3.140000104904175f));
    const v3 up = V(0.0f, 0.0f, 1.0f);
    float cvs = vdot(view, sd);
    float csu = vdot(sd, up);
    float cuv = vdot(up, view);
    if (sadc == 1.0f) return V(1.0f, 0.0f, 0.0f);
    float sunE = sun_intensity(csu);
    v3 ray = V(5.176821105123963e-6f, 1.2785348189936485e-5f, 2.8530756026157178e-5f);
    v3 mie = vscale(total_mie(V(6.800000278417428e-7f, 5.499999815583578e-7f, 4.4999998749517545e-7f),
                              V(0.6859999895095825f, 0.6779999732971191f, 0.6660000085830688f), 1.0f),
                    0.004999999888241291f);
    float zen = wmax(0.0f, cuv);
    float rol = 8400.0f / zen, mol = 1250.0f / zen;
    v3 fex = V(orc_det_expf(-((ray.x * rol) + (mie.x * mol))), orc_det_expf(-((ray.y * rol) + (mie.y * mol))),
               orc_det_expf(-((ray.z * rol) + (mie.z * mol))));
    float rp = rayleigh_phase(cvs), hp = hg_phase(cvs, 0.800000011920929f);
    v3 rxe = vscale(ray, rp), mxe = vscale(mie, hp);
    v3 tl = vadd(ray, mie), lx = vadd(rxe, mxe);
    v3 se = V(sunE * (lx.x / tl.x), sunE * (lx.y / tl.y), sunE * (lx.z / tl.z));
    v3 sky = V(se.x * (1.0f - fex.x), se.y * (1.0f - fex.y), se.z * (1.0f - fex.z));
    float f = wmin(wmax(orc_det_powf(1.0f - vdot(up, sd), 5.0f), 0.0f), 1.0f);
    v3 m = V(1.0f * (1.0f - f) + orc_det_powf(se.x * fex.x, 0.5f) * f,
             1.0f * (1.0f - f) + orc_det_powf(se.y * fex.y, 0.5f) * f,
             1.0f * (1.0f - f) + orc_det_powf(se.z * fex.z, 0.5f) * f);
    sky = vmul(sky, m);
    return vscale(sky, 0.009999999776482582f);
}

void orc_sunsky(const float view[3], float spx, float spy, float out[3])
{
    v3 r = sunsky(V(view[0], view[1], view[2]), spx, spy);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* LinearToSRGB (4600-4607), per channel: mix(pow(c, 1/2.4) * 1.055 - 0.055, c * 12.92, c < 0.0031308)
 * after clamping to [0, 1] */
float orc_linear_to_srgb(float x)
{
    float r = wmin(wmax(x, 0.0f), 1.0f);
    float a = orc_det_powf(r, 0.416666666666666667f) * 1.055f - 0.055f;
    float b = r * 12.92f;
    float t = r < 0.0031308f ? 1.0f : 0.0f;
    return a * (1.0f - t) + b * t;
}

typedef struct { float x, y, z, w; } v4;

/* get_color_single (compute.wgsl:4440-4523) for the scene-independent lighting (sunPos (-0.1,
 * 0.7), sun intensity 3, specular 1, albedo 1, sun size 0.01, background x1: the `scene`
 * branches of 4455-4468 are the reference's hand-written scenes).  The shadow ray's result is
 * read only when sunlight > 0 (4501), so it is traced only then.  *depth0: the first ray's
 * distance. */
static v4 color_single(const orc_state *s, v3 ro, v3 rd, float time, uint32_t *seed, uint64_t *visits,
                       float *depth0)
{
    const int cinematic = (s->p.shading & ORC_SHADE_CINEMATIC) != 0;
    const int normal_viz = (s->p.shading & ORC_SHADE_NORMALS) != 0;
    v3 frm = ro, dir = rd;
    v3 luminance = V(1.0f, 1.0f, 1.0f), direct = V(0.0f, 0.0f, 0.0f);
    const float spx = -0.1f, spy = 0.7f;
    const float sun_i = 3.0f, specular_i = 1.0f, albedo = 1.0f, sunsize = 0.01f, bgmul = 1.0f;
    int bounces = cinematic ? 2 : 1;
    for (int i = 0; i < bounces; i++) {
        v3 c, n;
        float d;
        trace_ray(s, frm, dir, time, &c, &d, visits, &n);
        if (i == 0 && depth0) *depth0 = d;
        v3 hitpoint = vadd(frm, vscale(dir, d));
        if (d >= 0.0f) {
            if (!cinematic) { v4 r = {c.x, c.y, c.z, 1.0f}; return r; }
            v3 old_rd = dir;
            dir = sample_biased(n, 3.0f, seed);
            luminance = vscale(vmul(luminance, c), albedo);
            frm = vadd(hitpoint, vscale(n, 0.05f));
            v3 sdn = vnormalize(from_spherical((spx - 0.0f) * 6.28f, (spy - 0.5f) * 3.14f));
            v3 sundir = V(sdn.x, sdn.z, sdn.y);
            float j0 = orc_frand(seed);
            float j1 = orc_frand(seed);
            float j2 = orc_frand(seed);
            v3 ssd = vnormalize(vadd(sundir, vscale(V(j0, j1, j2), sunsize)));
            float sunlight = vdot(n, ssd);
            if (sunlight > 0.0f) {
                v3 sc, sn;
                float sdist;
                trace_ray(s, frm, ssd, time, &sc, &sdist, visits, &sn);
                if (sdist < 0.0f) {
                    direct = vadd(direct, vscale(vscale(luminance, sunlight), sun_i));
                    v3 half = vnormalize(vadd(V(-old_rd.x, -old_rd.y, -old_rd.z), sundir));
                    float spec = (100.0f * orc_det_powf(wmax(vdot(n, half), 0.0f), 30.0f)) * specular_i;
                    direct = vadd(direct, vscale(luminance, spec));
                }
            }
        } else {
            if (i == 0) {                 /* see_background = false (3506) */
                if (cinematic || normal_viz) { v4 r = {10.0f, 10.0f, 10.0f, 0.0f}; return r; }
                v4 r = {c.x, c.y, c.z, 1.0f};
                return r;
            }
            v3 bg = sunsky(V(dir.x, dir.z, dir.y), spx, spy);
            v3 o = vadd(direct, vmul(luminance, bg));
            v4 r = {o.x * bgmul, o.y * bgmul, o.z * bgmul, 1.0f * bgmul};
            return r;
        }
    }
    v4 r = {direct.x, direct.y, direct.z, 1.0f};
    return r;
}

/* rt_main -> get_color_supersample with cinematic = true (4620-4648, 4651-4688): seed3 from the
 * pixel index and the time (4667), totalsamples = 8 rays jittered by frand() * 0.5, the mean
 * through LinearToSRGB, then the 0.45 gamma; alpha is the mean alpha.  depth: the first
 * sample's first ray. */
static int trace_cinematic(const orc_state *s, const float view[16], uint32_t W, uint32_t H, float time,
                           float zdist, float *rgba, float *depth, uint32_t row0, uint32_t r1,
                           uint64_t *node_visits)
{
    for (uint32_t y = row0; y < r1; y++) {
        for (uint32_t x = 0; x < W; x++) {
            uint32_t pi = y * W + x;
            uint32_t seed = pi + (uint32_t)(12345.0f * (orc_det_sinf(time * 12345.0f) * 0.5f + 0.5f));
            v3 eye = V(view[12], view[13], view[14]);
            v4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            float d0 = -1.0f;
            for (int k = 0; k < 8; k++) {
                float jx = orc_frand(&seed);
                float jy = orc_frand(&seed);
                float ox = (jx * 1.0f) * 0.5f, oy = (jy * 1.0f) * 0.5f;
                float px = (float)x + ox, py = (float)y + oy;
                /* ray_direction (4590-4594) */
                float xyx = (px - (float)W / 2.0f) + 0.5f;
                float xyy = (py - (float)H / 2.0f) + 0.5f;
                v3 d = V(xyx, -xyy, -zdist);
                float len = vlength(d);
                d = V(d.x / len, d.y / len, d.z / len);
                v3 g = V(((view[0] * d.x + view[4] * d.y) + view[8] * d.z) + view[12] * 0.0f,
                         ((view[1] * d.x + view[5] * d.y) + view[9] * d.z) + view[13] * 0.0f,
                         ((view[2] * d.x + view[6] * d.y) + view[10] * d.z) + view[14] * 0.0f);
                v4 c = color_single(s, eye, g, time, &seed, node_visits, k == 0 ? &d0 : NULL);
                acc.x = acc.x + c.x;
                acc.y = acc.y + c.y;
                acc.z = acc.z + c.z;
                acc.w = acc.w + c.w;
            }
            acc.x = acc.x / 8.0f;
            acc.y = acc.y / 8.0f;
            acc.z = acc.z / 8.0f;
            acc.w = acc.w / 8.0f;
            size_t o = (size_t)pi;
            if (rgba) {
                rgba[4 * o + 0] = orc_det_pow045(orc_linear_to_srgb(acc.x));
                rgba[4 * o + 1] = orc_det_pow045(orc_linear_to_srgb(acc.y));
                rgba[4 * o + 2] = orc_det_pow045(orc_linear_to_srgb(acc.z));
                rgba[4 * o + 3] = acc.w;
            }
            if (depth) depth[o] = d0;
        }
    }
    return ORC_OK;
}

int orc_trace(const orc_state *s, const float view[16], uint32_t W, uint32_t H, float time, float zdist,
              float *rgba, float *depth, uint32_t row0, uint32_t rows, uint64_t *node_visits)
{
    uint32_t r1 = rows ? row0 + rows : H;
    if (r1 > H) r1 = H;
    if (s->p.shading & ORC_SHADE_CINEMATIC)
        return trace_cinematic(s, view, W, H, time, zdist, rgba, depth, row0, r1, node_visits);
    for (uint32_t y = row0; y < r1; y++) {
        for (uint32_t x = 0; x < W; x++) {
            /* ray_direction, compute.wgsl:4590-4594 */
            float xyx = ((float)x - (float)W / 2.0f) + 0.5f;
            float xyy = ((float)y - (float)H / 2.0f) + 0.5f;
            v3 d = V(xyx, -xyy, -zdist);
            float len = vlength(d);
            d = V(d.x / len, d.y / len, d.z / len);
            /* view * vec4(d, 0), compute.wgsl:4630 */
            v3 g = V(((view[0] * d.x + view[4] * d.y) + view[8] * d.z) + view[12] * 0.0f,
                     ((view[1] * d.x + view[5] * d.y) + view[9] * d.z) + view[13] * 0.0f,
                     ((view[2] * d.x + view[6] * d.y) + view[10] * d.z) + view[14] * 0.0f);
            v3 eye = V(view[12], view[13], view[14]);
            v3 c;
            float dist;
            trace_ray(s, eye, g, time, &c, &dist, node_visits, NULL);
            /* get_color_single returns vec4(c, 1) on hit and on the first-bounce miss, and
             * vec4(10, 10, 10, 0) on a miss with normal_viz (4510-4514);
             * get_color_supersample: c = (0 + c) / 1, then pow(c, 0.45) per channel. */
            float ca = 1.0f;
            if ((s->p.shading & ORC_SHADE_NORMALS) && !(dist >= 0.0f)) {
                c = V(10.0f, 10.0f, 10.0f);
                ca = 0.0f;
            }
            float r = (0.0f + c.x) / 1.0f, gg = (0.0f + c.y) / 1.0f, b = (0.0f + c.z) / 1.0f, a = (0.0f + ca) / 1.0f;
            size_t pi = (size_t)y * W + x;
            if (rgba) {
                rgba[4 * pi + 0] = orc_det_pow045(r);
                rgba[4 * pi + 1] = orc_det_pow045(gg);
                rgba[4 * pi + 2] = orc_det_pow045(b);
                rgba[4 * pi + 3] = a;
            }
            if (depth) depth[pi] = dist;
        }
    }
    return ORC_OK;
}

/* ---------------- canonical export ---------------- */

typedef struct { uint32_t v, d, i, j, k; } frame_t;

uint64_t orc_export(const orc_state *s, orc_node *out, uint64_t cap)
{
    uint64_t cnt = 0, sp = 0, scap = 1024;
    frame_t *st = (frame_t *)malloc(scap * sizeof(frame_t));
    if (!st) return 0;
    frame_t f0 = {0, 0, 0, 0, 0};
    st[sp++] = f0;
    while (sp) {
        frame_t f = st[--sp];
        const orc_voxel *v = &s->vox[f.v];
        if (out && cnt < cap) {
            orc_node *o = &out[cnt];
            o->depth = f.d; o->i = f.i; o->j = f.j; o->k = f.k;
            o->f = v->f; o->t = v->t; o->time_err = v->time_err; o->fast = v->fast;
            o->has_children = v->child_count != 0u;
            o->dx = v->dx; o->dy = v->dy; o->dz = v->dz;
        }
        cnt++;
        if (v->child_count != 0u) {
            for (uint32_t l = 0; l < 8; l++) {
                if (sp == scap) {
                    scap *= 2;
                    frame_t *ns = (frame_t *)realloc(st, scap * sizeof(frame_t));
                    if (!ns) { free(st); return cnt; }
                    st = ns;
                }
                frame_t c = {v->children[l], f.d + 1, 2 * f.i + (l & 1u), 2 * f.j + ((l >> 1) & 1u), 2 * f.k + (l >> 2)};
                st[sp++] = c;
            }
        }
    }
    free(st);
    return cnt;
}

const orc_voxel *orc_voxels(const orc_state *s, uint32_t *n_used)
{
    if (n_used) *n_used = s->start_of_empty_voxel + 1;
    return s->vox;
}
