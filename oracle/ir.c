/*
 * ir.c -- tape builder of the oracle: restatement of src/ir.rs (the .ll -> tape builder).
 * TEST INFRASTRUCTURE ONLY (see tsi_oracle.h).
 *
 * Mirrors, in order:
 *   ir.rs:8-43     opcode numbering
 *   ir.rs:45-58    UNARY_OPS / TAIL_OPS membership
 *   ir.rs:190-331  parseLLVMIR: locate "fffff", walk the lines of the body up to the
 *                  first '}', classify by token, parse operands, reindex %k>3 -> k-1,
 *                  MUL x,x -> SQR, count uses per register operand
 *   ir.rs:334-366  IR::new: four ARG instructions, box [-5,5]^3 and t, dummy use on RET
 *   ir.rs:139-181  Instruction::to_gpu: literal side bits, pack2x16 packing
 * Where ir.rs would panic (an unwrap on a malformed token) this returns ORC_E_PARSE.
 */
#include "tsi_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <errno.h>

enum { OP_NOP = 0, OP_ARGX, OP_ARGY, OP_ARGZ, OP_ARGW, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_EXP, OP_LOG,
       OP_SQR, OP_SQRT, OP_POW, OP_SIN, OP_COS, OP_NEG, OP_MIN, OP_MAX, OP_ABS, OP_CUP, OP_CAP, OP_BR,
       OP_PHI, OP_TRUNC, OP_FLOOR, OP_CEIL, OP_SIGN, OP_RET, OP_MOD, OP_ATAN2 };

/* ir.rs:47-50 */
static int is_unary(int op)
{
    switch (op) {
    case OP_NOP: case OP_ARGX: case OP_ARGY: case OP_ARGZ: case OP_LOG: case OP_SQR: case OP_SQRT:
    case OP_SIN: case OP_COS: case OP_NEG: case OP_EXP: case OP_TRUNC: case OP_FLOOR: case OP_CEIL:
    case OP_SIGN: case OP_ABS: case OP_RET:
        return 1;
    default:
        return 0;
    }
}

/* ir.rs:54 */
static int is_tail(int op)
{
    switch (op) {
    case OP_EXP: case OP_LOG: case OP_SQRT: case OP_MIN: case OP_MAX: case OP_FLOOR: case OP_CEIL:
    case OP_ABS: case OP_SIN: case OP_COS: case OP_POW: case OP_MOD: case OP_ATAN2:
        return 1;
    default:
        return 0;
    }
}

typedef struct { int is_reg; uint32_t reg; float lit; } arg_t;   /* ir.rs:68-71 ARG */

typedef struct {
    uint32_t index, op, uses;
    arg_t larg, rarg;
    float val[2];
} ins_t;

#define MAXTOK 24
typedef struct { const char *p; size_t n; } tok_t;

static int split_ws(const char *line, size_t len, tok_t *tok, int max)
{
    int nt = 0;
    size_t i = 0;
    while (i < len) {
        while (i < len && (line[i] == ' ' || line[i] == '\t' || line[i] == '\r' || line[i] == '\v' || line[i] == '\f'))
            i++;
        if (i >= len)
            break;
        size_t s = i;
        while (i < len && !(line[i] == ' ' || line[i] == '\t' || line[i] == '\r' || line[i] == '\v' || line[i] == '\f'))
            i++;
        if (nt < max) {
            tok[nt].p = line + s;
            tok[nt].n = i - s;
        }
        nt++;
    }
    return nt;
}

static int contains(tok_t t, const char *needle)
{
    size_t m = strlen(needle);
    if (m > t.n)
        return 0;
    for (size_t i = 0; i + m <= t.n; i++)
        if (memcmp(t.p + i, needle, m) == 0)
            return 1;
    return 0;
}

static int teq(tok_t t, const char *s) { return t.n == strlen(s) && memcmp(t.p, s, t.n) == 0; }

/* Rust str::parse::<u32>() on t[a..t.n-b] */
static int parse_u32(tok_t t, size_t a, size_t b, uint32_t *out)
{
    if (a + b > t.n || a + b == t.n)
        return 0;
    uint64_t v = 0;
    size_t i = a, e = t.n - b;
    if (t.p[i] == '+')
        i++;
    if (i >= e)
        return 0;
    for (; i < e; i++) {
        if (t.p[i] < '0' || t.p[i] > '9')
            return 0;
        v = v * 10 + (uint64_t)(t.p[i] - '0');
        if (v > 0xFFFFFFFFu)
            return 0;
    }
    *out = (uint32_t)v;
    return 1;
}

/* Rust str::parse::<f32>() (correctly rounded decimal) on t[a..t.n-b] */
static int parse_f32(tok_t t, size_t a, size_t b, float *out)
{
    if (a + b >= t.n)
        return 0;
    char buf[128];
    size_t m = t.n - a - b;
    if (m >= sizeof buf)
        return 0;
    memcpy(buf, t.p + a, m);
    buf[m] = 0;
    for (size_t i = 0; i < m; i++) {
        char c = buf[i];
        if (!((c >= '0' && c <= '9') || c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-'))
            return 0;
    }
    char *end = NULL;
    errno = 0;
    float v = strtof(buf, &end);
    if (end != buf + m)
        return 0;
    *out = v;
    return 1;
}

/* u64::from_str_radix(t[a..t.n-b], 16) -> f64::from_bits -> as f32 */
static int parse_hex(tok_t t, size_t a, size_t b, float *out)
{
    if (a + b >= t.n)
        return 0;
    uint64_t v = 0;
    size_t e = t.n - b;
    if (e - a > 16)
        return 0;
    for (size_t i = a; i < e; i++) {
        char c = t.p[i];
        int d;
        if (c >= '0' && c <= '9') d = c - '0';
        else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
        else if (c >= 'A' && c <= 'F') d = c - 'A' + 10;
        else return 0;
        v = (v << 4) | (uint64_t)d;
    }
    double dv;
    memcpy(&dv, &v, 8);
    *out = (float)dv;
    return 1;
}

static int arg_eq(arg_t a, arg_t b)   /* #[derive(PartialEq)] on ARG */
{
    if (a.is_reg != b.is_reg)
        return 0;
    return a.is_reg ? a.reg == b.reg : a.lit == b.lit;
}

static uint32_t pack2x16(uint32_t a, uint32_t b) { return ((a & 0xFFFFu) << 16) | (b & 0xFFFFu); }   /* ir.rs:99 */

/* ir.rs:140-181 */
static orc_instr to_gpu(const ins_t *in)
{
    orc_instr g;
    float literal = 0.0f;
    uint32_t side = 3;
    uint32_t larg = 0, rarg = 0;
    if (in->larg.is_reg) { side -= 2; larg = in->larg.reg & 0xFFFFu; }
    else literal = in->larg.lit;
    if (in->rarg.is_reg) { side -= 1; rarg = in->rarg.reg & 0xFFFFu; }
    else literal = in->rarg.lit;
    uint32_t op = (in->op | (side << 14)) & 0xFFFFu;
    g.index_ = in->index & 0xFFFFu;
    g.opcode_uses = pack2x16(op, in->uses);
    g.larg_rarg = pack2x16(larg, rarg);
    g.literal = literal;
    g.val[0] = in->val[0];
    g.val[1] = in->val[1];
    g.sens[0] = 0.0f;
    g.sens[1] = 0.0f;
    return g;
}

static arg_t REG(uint32_t r) { arg_t a = {1, r, 0.0f}; return a; }
static arg_t LIT(float v) { arg_t a = {0, 0, v}; return a; }

int orc_parse_ll(const char *text, size_t len, float time, orc_instr *out, uint32_t cap, uint32_t *n_total)
{
    /* ir.rs:191-196 */
    const char *a = NULL;
    for (size_t i = 0; i + 5 <= len; i++)
        if (memcmp(text + i, "fffff", 5) == 0) { a = text + i; break; }
    if (!a)
        return ORC_E_PARSE;
    size_t blen = len - (size_t)(a - text);
    const char *nl = memchr(a, '\n', blen);
    const char *br = memchr(a, '}', blen);
    if (!nl || !br || br <= nl)
        return ORC_E_PARSE;
    const char *fbeg = nl, *fend = br - 1;   /* f = &b[start..end-1] */

    size_t ninst = 0, capi = 1024;
    ins_t *ins = (ins_t *)calloc(capi, sizeof(ins_t));
    if (!ins)
        return ORC_E_PARSE;
    /* ir.rs:334-348 */
    for (uint32_t k = 0; k < 4; k++) {
        ins[k].index = k;
        ins[k].op = OP_ARGX + k;
        ins[k].larg = LIT(0.0f);
        ins[k].rarg = LIT(0.0f);
    }
    ins[0].val[0] = ins[1].val[0] = ins[2].val[0] = -5.0f;
    ins[0].val[1] = ins[1].val[1] = ins[2].val[1] = 5.0f;
    ins[3].val[0] = ins[3].val[1] = time;
    ninst = 4;

    int rc = ORC_OK;
    /* lines = f.split("\n"); for i in 1..lines.len() */
    const char *p = fbeg;
    int first = 1;
    while (p <= fend) {
        const char *q = memchr(p, '\n', (size_t)(fend - p));
        const char *lend = q ? q : fend;
        if (first) { first = 0; p = lend + 1; if (!q) break; continue; }
        tok_t tk[MAXTOK];
        int nt = split_ws(p, (size_t)(lend - p), tk, MAXTOK);
        p = lend + 1;
        if (nt > MAXTOK) nt = MAXTOK;
        if (ninst + 1 > capi) {
            capi *= 2;
            ins_t *g = (ins_t *)realloc(ins, capi * sizeof(ins_t));
            if (!g) { rc = ORC_E_PARSE; break; }
            ins = g;
        }
        ins_t in;
        memset(&in, 0, sizeof in);
        in.index = (uint32_t)ninst & 0xFFFFu;
        in.larg = LIT(0.0f);
        in.rarg = LIT(0.0f);
        if (nt < 1) { rc = ORC_E_PARSE; break; }
        /* ir.rs:204-235 */
        if (teq(tk[0], "ret")) {
            in.op = OP_RET;
        } else {
            if (nt < 3) { rc = ORC_E_PARSE; break; }
            if (teq(tk[2], "fadd")) in.op = OP_ADD;
            else if (teq(tk[2], "fneg")) in.op = OP_NEG;
            else if (teq(tk[2], "fsub")) in.op = OP_SUB;
            else if (teq(tk[2], "fmul")) in.op = OP_MUL;
            else if (teq(tk[2], "fdiv")) in.op = OP_DIV;
            else if (teq(tk[2], "tail")) {
                if (nt < 6) { rc = ORC_E_PARSE; break; }
                tok_t t5 = tk[5];
                if (contains(t5, "sqrt")) in.op = OP_SQRT;
                else if (contains(t5, "log")) in.op = OP_LOG;
                else if (contains(t5, "exp")) in.op = OP_EXP;
                else if (contains(t5, "min")) in.op = OP_MIN;
                else if (contains(t5, "max")) in.op = OP_MAX;
                else if (contains(t5, "floor")) in.op = OP_FLOOR;
                else if (contains(t5, "abs")) in.op = OP_ABS;
                else if (contains(t5, "sin")) in.op = OP_SIN;
                else if (contains(t5, "cos")) in.op = OP_COS;
                else if (contains(t5, "pow")) in.op = OP_POW;
                else if (contains(t5, "mod")) in.op = OP_MOD;
                else if (contains(t5, "atan2")) in.op = OP_ATAN2;
                else in.op = OP_NOP;   /* "cannot parse" */
            } else {
                in.op = OP_NOP;        /* "cannot parse" */
            }
        }
        int op = (int)in.op;
        int ok = 1;
        if (is_tail(op) && is_unary(op)) {                   /* ir.rs:238-247 */
            if (nt < 7) { rc = ORC_E_PARSE; break; }
            tok_t t = tk[6];
            float v;
            uint32_t r;
            if (contains(t, "%")) { ok = parse_u32(t, 1, 1, &r); in.larg = REG(r); }
            else if (contains(t, "0x")) { ok = parse_hex(t, 2, 1, &v); in.larg = LIT(v); }
            else { ok = parse_f32(t, 0, 0, &v); in.larg = LIT(v); }
        } else if (is_tail(op) && !is_unary(op)) {           /* ir.rs:248-262 */
            if (nt < 9) { rc = ORC_E_PARSE; break; }
            tok_t t = tk[6];
            float v;
            uint32_t r;
            if (contains(t, "%")) { ok = parse_u32(t, 1, 1, &r); in.larg = REG(r); }
            else { ok = parse_f32(t, 0, 0, &v); in.larg = LIT(v); }
            t = tk[8];
            if (ok) {
                if (contains(t, "%")) { ok = parse_u32(t, 1, 1, &r); in.rarg = REG(r); }
                else if (contains(t, "0x")) { ok = parse_hex(t, 2, 1, &v); in.rarg = LIT(v); }
                else { ok = parse_f32(t, 0, 1, &v); in.rarg = LIT(v); }
            }
        } else if (is_unary(op)) {                           /* ir.rs:263-272 */
            uint32_t r;
            float v;
            if (teq(tk[0], "ret")) {
                if (nt < 3) { rc = ORC_E_PARSE; break; }
                ok = parse_u32(tk[2], 1, 0, &r);
                in.larg = REG(r);
            } else {
                if (nt < 5) { rc = ORC_E_PARSE; break; }
                tok_t t = tk[4];
                if (contains(t, "%")) { ok = parse_u32(t, 1, 0, &r); in.larg = REG(r); }
                else { ok = parse_f32(t, 0, 0, &v); in.larg = LIT(v); }
            }
            in.rarg = LIT(0.0f);
        } else {                                             /* ir.rs:273-293 */
            if (nt < 6) { rc = ORC_E_PARSE; break; }
            tok_t t = tk[4];
            float v;
            uint32_t r;
            if (contains(t, "%")) { ok = parse_u32(t, 1, 1, &r); in.larg = REG(r); }
            else if (contains(t, "0x")) { ok = parse_hex(t, 2, 1, &v); in.larg = LIT(v); }
            else { ok = parse_f32(t, 0, 1, &v); in.larg = LIT(v); }
            t = tk[5];
            if (ok) {
                if (contains(t, "%")) { ok = parse_u32(t, 1, 0, &r); in.rarg = REG(r); }
                else if (contains(t, "0x")) { ok = parse_hex(t, 2, 0, &v); in.rarg = LIT(v); }
                else { ok = parse_f32(t, 0, 0, &v); in.rarg = LIT(v); }
            }
        }
        if (!ok) { rc = ORC_E_PARSE; break; }
        /* ir.rs:297-307 reindexing */
        if (in.larg.is_reg && in.larg.reg > 3) in.larg.reg -= 1;
        if (in.rarg.is_reg && in.rarg.reg > 3) in.rarg.reg -= 1;
        /* ir.rs:310-313 */
        if (in.op == OP_MUL && arg_eq(in.larg, in.rarg)) {
            in.op = OP_SQR;
            in.rarg = LIT(0.0f);
        }
        /* ir.rs:315-327 */
        if (in.larg.is_reg) {
            if (in.larg.reg >= ninst) { rc = ORC_E_PARSE; break; }
            ins[in.larg.reg].uses++;
        }
        if (in.rarg.is_reg) {
            if (in.rarg.reg >= ninst) { rc = ORC_E_PARSE; break; }
            ins[in.rarg.reg].uses++;
        }
        ins[ninst++] = in;
        if (!q) break;
    }
    if (rc == ORC_OK) {
        ins[ninst - 1].uses++;   /* ir.rs:355 dummy use on the last instruction */
        if (n_total) *n_total = (uint32_t)ninst;
        uint32_t m = (uint32_t)(ninst < 4096 ? ninst : 4096);
        if (m > cap) m = cap;
        for (uint32_t i = 0; i < m; i++)
            out[i] = to_gpu(&ins[i]);
        if (ninst > 4096) rc = ORC_E_TAPE_TOO_LONG;
    }
    free(ins);
    return rc;
}
