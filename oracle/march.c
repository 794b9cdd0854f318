/*
 * march.c -- float64 ray march of an octree leaf set.  TEST INFRASTRUCTURE ONLY (see
 * tsi_oracle.h): the independent check of the depth the trace reports.
 *
 * The trace of the reference (compute.wgsl:4189-4403) walks the octree with a 2x2x2 DDA in
 * binary32 and returns |p_hit - eye| for the first leaf with f <= 0 whose box, widened by 1e-4
 * (4310-4311), the ray enters; it descends only into children with f <= 0 (4302).  Here the
 * same question is answered without a DDA: front-to-back recursion over exact slab tests in
 * float64 (children visited in order of their entry distance, subtrees farther than the best
 * hit so far skipped).  The nodes are the canonical (depth, i, j, k) export of orc_export /
 * tsi_read_nodes; a node's box is the root box cut into 2^depth cells per axis.
 */
#include "tsi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc_leafset {
    uint64_t *keys;          /* open addressing, 0 = empty (key + 1 stored) */
    uint32_t *idx;
    uint64_t mask;
    const orc_node *nodes;   /* borrowed for the lifetime of the set */
    double lo[3], hi[3], pad;
};

static uint64_t node_key(uint32_t d, uint32_t i, uint32_t j, uint32_t k)
{
    return ((uint64_t)d << 60) | ((uint64_t)i << 40) | ((uint64_t)j << 20) | (uint64_t)k;
}

static uint64_t mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

static int64_t find(const orc_leafset *s, uint64_t key)
{
    uint64_t h = mix64(key) & s->mask;
    for (;;) {
        if (s->keys[h] == 0) return -1;
        if (s->keys[h] == key + 1) return s->idx[h];
        h = (h + 1) & s->mask;
    }
}

orc_leafset *orc_leafset_build(const orc_node *nodes, uint64_t n, const double root_box[6], double pad)
{
    if (!nodes || n == 0 || n > 0xFFFFFFFFull) return NULL;
    orc_leafset *s = (orc_leafset *)calloc(1, sizeof *s);
    if (!s) return NULL;
    uint64_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    s->keys = (uint64_t *)calloc(cap, sizeof(uint64_t));
    s->idx = (uint32_t *)malloc(cap * sizeof(uint32_t));
    if (!s->keys || !s->idx) { orc_leafset_free(s); return NULL; }
    s->mask = cap - 1;
    s->nodes = nodes;
    for (int a = 0; a < 3; a++) { s->lo[a] = root_box[a]; s->hi[a] = root_box[3 + a]; }
    s->pad = pad;
    for (uint64_t e = 0; e < n; e++) {
        const orc_node *o = &nodes[e];
        if (o->depth > 15 || o->i >= (1u << 20) || o->j >= (1u << 20) || o->k >= (1u << 20)) {
            orc_leafset_free(s);
            return NULL;
        }
        uint64_t key = node_key(o->depth, o->i, o->j, o->k);
        uint64_t h = mix64(key) & s->mask;
        while (s->keys[h] != 0 && s->keys[h] != key + 1) h = (h + 1) & s->mask;
        s->keys[h] = key + 1;
        s->idx[h] = (uint32_t)e;
    }
    return s;
}

void orc_leafset_free(orc_leafset *s)
{
    if (!s) return;
    free(s->keys);
    free(s->idx);
    free(s);
}

/* entry distance of the ray into [lo, hi] (clamped at 0), or +inf if it misses */
static double slab(const double lo[3], const double hi[3], const double o[3], const double inv[3])
{
    double t0 = 0.0, t1 = INFINITY;
    for (int a = 0; a < 3; a++) {
        double u = (lo[a] - o[a]) * inv[a], v = (hi[a] - o[a]) * inv[a];
        if (u != u) u = -INFINITY;   /* ray in the slab's plane: no constraint from this axis */
        if (v != v) v = INFINITY;
        double mn = u < v ? u : v, mx = u < v ? v : u;
        if (mn > t0) t0 = mn;
        if (mx < t1) t1 = mx;
    }
    return t0 <= t1 ? t0 : INFINITY;
}

static void cell_box(const orc_leafset *s, uint32_t d, uint32_t i, uint32_t j, uint32_t k, double pad, double lo[3],
                     double hi[3])
{
    const uint32_t c[3] = {i, j, k};
    for (int a = 0; a < 3; a++) {
        double w = (s->hi[a] - s->lo[a]) / (double)(1u << d);
        lo[a] = s->lo[a] + w * c[a] - pad;
        hi[a] = s->lo[a] + w * (c[a] + 1) + pad;
    }
}

/* first hit under node e (its box already entered), or best if none closer */
static double visit(const orc_leafset *s, uint32_t e, const double o[3], const double inv[3], double best)
{
    const orc_node *p = &s->nodes[e];
    double tc[8];
    int64_t ci[8];
    int order[8], m = 0;
    for (uint32_t l = 0; l < 8; l++) {
        uint32_t d = p->depth + 1, i = 2 * p->i + (l & 1u), j = 2 * p->j + ((l >> 1) & 1u), k = 2 * p->k + (l >> 2);
        int64_t c = find(s, node_key(d, i, j, k));
        if (c < 0 || !(s->nodes[c].f <= 0.0f)) continue;   /* compute.wgsl:4302 */
        double lo[3], hi[3];
        cell_box(s, d, i, j, k, s->pad, lo, hi);
        double t = slab(lo, hi, o, inv);
        if (!(t < best)) continue;
        tc[m] = t;
        ci[m] = c;
        order[m] = m;
        m++;
    }
    for (int a = 1; a < m; a++)   /* front to back */
        for (int b = a; b > 0 && tc[order[b]] < tc[order[b - 1]]; b--) {
            int x = order[b]; order[b] = order[b - 1]; order[b - 1] = x;
        }
    for (int a = 0; a < m; a++) {
        int q = order[a];
        if (!(tc[q] < best)) break;
        const orc_node *c = &s->nodes[ci[q]];
        if (!c->has_children) best = tc[q];
        else best = visit(s, (uint32_t)ci[q], o, inv, best);
    }
    return best;
}

void orc_leafset_march(const orc_leafset *s, const double *origins, const double *dirs, uint64_t m, double *depth)
{
    int64_t root = find(s, 0);
    for (uint64_t r = 0; r < m; r++) {
        const double *o = origins + 3 * r, *dv = dirs + 3 * r;
        double inv[3] = {1.0 / dv[0], 1.0 / dv[1], 1.0 / dv[2]};
        double lo[3], hi[3];
        cell_box(s, 0, 0, 0, 0, 0.0, lo, hi);   /* the root box is not widened (4206-4209) */
        double best = INFINITY;
        if (root >= 0 && s->nodes[root].has_children && slab(lo, hi, o, inv) < INFINITY)
            best = visit(s, (uint32_t)root, o, inv, INFINITY);
        depth[r] = best < INFINITY ? best : -1.0;
    }
}
