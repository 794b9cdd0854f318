// tsi_jit.cpp -- per-tape straight-line evaluation kernels, compiled at tape upload with hiprtc.
//
// The reference interprets the tape in every lane (compute.wgsl:3633-3964: a switch per
// instruction over values held in global memory).  A tape is fixed for the lifetime of a
// scene, so here it is also turned into code: every record becomes the interval rule of its
// form applied to named values (tsi_rules.h, the same functions the interpreter calls), the
// reverse pass becomes the j-descending accumulation written out with the first-touch and
// short-circuit decisions resolved at generation time, and the result is compiled for gfx950
// into the level kernel of tsi_eval.h (TapeEngine).  Values then live in registers allocated
// by the compiler, records cost no fetch or dispatch, and independent records overlap.
//
// Saved operand values go through the same per-wavefront stack as in the interpreter
// (positions from the lowering), which keeps the live ranges of the forward values short.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <sys/stat.h>
#include <unistd.h>
#include <utime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <iterator>
#include <vector>

#include "tsi.h"
#include "tsi_jit.h"
#include "tsi_tape.h"

extern "C" {
extern const int tsi_jit_num_headers;
extern const char *const tsi_jit_header_names[];
extern const char *const tsi_jit_header_texts[];
}

namespace tsi {

namespace {

inline uint32_t op_of(const tsi_gpu_instruction &g) { return (g.opcode_uses >> 16) & ~(3u << 14); }
inline uint32_t side_of(const tsi_gpu_instruction &g) { return (g.opcode_uses >> 16) >> 14; }
inline uint32_t larg_of(const tsi_gpu_instruction &g) { return g.larg_rarg >> 16; }
inline uint32_t rarg_of(const tsi_gpu_instruction &g) { return g.larg_rarg & 0xFFFFu; }

std::string lit(float c) {
    uint32_t u;
    std::memcpy(&u, &c, 4);
    char b[48];
    std::snprintf(b, sizeof b, "__uint_as_float(0x%08xu)", u);
    return b;
}

// Records per region ($TSI_JIT_REGION, else 16 in pruned modules, whose region bits are the
// pruning unit, and 64 in unpruned ones, where a longer region hides more of the next region's
// load latency and leaves the scheduler longer blocks: on the 1024^3 CSG frame 24, 32, 48, 64,
// 96 and 128 records gave 381, 384-385, 386, 388-389, 389 and 391 M cell-evaluations/s).  Each region is its own basic block, entered through a uniform branch on
// a value the compiler cannot see through (one scalar instruction), so that instruction
// selection and scheduling stay local: without it the whole tape is one block in which every
// load is hoisted to the top and every operation sunk to the bottom.  The reverse pass issues
// the saved-operand loads of the next region in the current one.
uint32_t region_records(bool prune) {
    static const uint32_t g = [] {
        const char *e = std::getenv("TSI_JIT_REGION");
        const long v = e ? std::strtol(e, nullptr, 10) : 0;
        return v >= 4 && v <= 1024 ? uint32_t(v) : 0u;
    }();
    return g ? g : (prune ? 16u : 64u);
}

// Regions ahead whose saved-operand loads the reverse pass of an unpruned module has in flight
// ($TSI_JIT_PREFETCH, default 1: the next region's loads are issued before this region runs).
uint32_t prefetch_regions() {
    static const uint32_t d = [] {
        const char *e = std::getenv("TSI_JIT_PREFETCH");
        const long v = e ? std::strtol(e, nullptr, 10) : 0;
        return v >= 1 && v <= 8 ? uint32_t(v) : 1u;
    }();
    return d;
}

// Minimum wavefronts per SIMD the level kernel is compiled for ($TSI_JIT_WAVES, 0: the
// compiler's choice).  Default 4, i.e. at most 128 VGPRs: the 3 685-record CSG module then
// spills 4 VGPRs (it held 167, 3 wavefronts per SIMD) and its frames run 4-5 % faster; 5
// wavefronts (96 VGPRs) spill 41 for no further gain.
uint32_t level_waves() {
    static const uint32_t w = [] {
        const char *e = std::getenv("TSI_JIT_WAVES");
        const long v = e ? std::strtol(e, nullptr, 10) : 4;
        return v >= 1 && v <= 8 ? uint32_t(v) : 0u;
    }();
    return w;
}

// Largest cost (in estimated VALU instructions per interval) of the forward code that the reverse
// pass of an unpruned module reruns to recompute a value instead of saving it
// ($TSI_JIT_REMAT, default 160; 0 keeps the one-step recomputation of literal forms only).  On the
// 1024^3 CSG frame (343 saved values without it), budgets of 80, 160 and 320 save 168, 74 and
// fewer and gave 334, 355 and 353 M cell-evaluations/s (305 without).
uint32_t remat_budget() {
    static const uint32_t b = [] {
        const char *e = std::getenv("TSI_JIT_REMAT");
        const long v = e ? std::strtol(e, nullptr, 10) : 160;
        return v >= 0 && v <= 100000 ? uint32_t(v) : 160u;
    }();
    return b;
}

// Estimated VALU instructions of a form's forward value (0: never recomputed: the transcendental
// and division forms, whose code is long).
uint32_t recompute_cost(uint32_t form) {
    switch (form) {
    case F_ADD_RR: case F_SUB_RR: case F_ADD_RL: case F_SUB_RL: case F_SUB_LR: case F_NEG: case F_MUL_RL: return 2;
    case F_MIN_RR: case F_MAX_RR: case F_MIN_RL: case F_MAX_RL: return 4;
    case F_ABS: return 6;
    case F_SQR: return 8;
    case F_MUL_RR: return 12;
    case F_SQRT: return 36;
    default: return 0;
    }
}

// Wavefront priority of the level kernel's frame levels (s_setprio; $TSI_JIT_PRIO, 0-3, default
// 3): the previous frame's trace runs beside them on the same SIMDs and is not on the frame's
// critical path.  On the 1024^3 CSG frame: 376-377 M cell-evaluations/s at 0, 382 at 3.
uint32_t level_prio() {
    static const uint32_t p = [] {
        const char *e = std::getenv("TSI_JIT_PRIO");
        const long v = e ? std::strtol(e, nullptr, 10) : 3;
        return v >= 0 && v <= 3 ? uint32_t(v) : 3u;
    }();
    return p;
}

// Records per region of batch modules ($TSI_JIT_BATCH_REGION, default 32).
uint32_t batch_region_records() {
    static const uint32_t g = [] {
        const char *e = std::getenv("TSI_JIT_BATCH_REGION");
        const long v = e ? std::strtol(e, nullptr, 10) : 32;
        return v >= 4 && v <= 1024 ? uint32_t(v) : 32u;
    }();
    return g;
}

// Batch modules with the skipped subtrees ($TSI_JIT_BATCH_SKIP=1; default without).
bool batch_skip() {
    static const bool b = [] {
        const char *e = std::getenv("TSI_JIT_BATCH_SKIP");
        return e && std::strtol(e, nullptr, 10) != 0;
    }();
    return b;
}

// Smallest subtree (records dominated by one record) that the reverse pass of an unpruned module
// skips when the dominating record's sensitivity is [0, 0] in every lane of the wavefront
// ($TSI_JIT_SKIP, default 8; 0 turns the skipping off).  Each guard costs a uniform test and a
// block boundary: on the 1024^3 CSG frame, minimums of 1, 4, 8 and 16 gave 287, 302, 311 and
// 300 M cell-evaluations/s (268 without skipping).
uint32_t skip_min_records() {
    static const uint32_t m = [] {
        const char *e = std::getenv("TSI_JIT_SKIP");
        const long v = e ? std::strtol(e, nullptr, 10) : 8;
        return v >= 0 && v <= 100000 ? uint32_t(v) : 8u;
    }();
    return m;
}

std::string V(uint32_t j) { return "v" + std::to_string(j); }
std::string S(uint32_t j) { return "s" + std::to_string(j); }
// row pos of the wavefront's stack, this lane's column (uniform row base + lane offset)
std::string P(uint32_t pos) { return "(stk + " + std::to_string(size_t(pos) * 64u) + "u)[lane]"; }

const char *kOpen = "    {\n        uint32_t go;\n        asm volatile(\"s_mov_b32 %0, 1\" : \"=s\"(go));\n        if (go) {\n";
const char *kClose = "        }\n    }\n";

}  // namespace

bool is_minmax_form(uint32_t form) {
    return form == F_MIN_RR || form == F_MAX_RR || form == F_MIN_RL || form == F_MAX_RL;
}

// forms whose reverse rule needs only a decision (compute.wgsl 2640 ABS_I_RR, 2884-2903 MIN/MAX)
bool is_code_form(uint32_t form) { return is_minmax_form(form) || form == F_ABS; }

// operand values the reverse rule of a form reads (rev_form_t in tsi_rules.h): x = the first
// register operand (the register of a literal-then-register form), y = the second
void rev_needs(uint32_t form, bool &x, bool &y) {
    switch (form) {
    case F_MUL_RR: case F_DIV_RR: case F_POW_RR: case F_ATAN2_RR: x = y = true; return;
    case F_SQR: case F_SQRT: case F_SIN: case F_COS: case F_EXP: case F_LOG: case F_FLOOR:
    case F_POW_RL: case F_ATAN2_RL: case F_DIV_LR:
        x = true; y = false; return;
    default: x = y = false; return;
    }
}

// ---- where fix() can be left out ----
// The binary32 interval rules of the forms below are inclusion-monotone: rounding to nearest is
// monotone, so a rule applied to sub-intervals of its operands returns a sub-interval of what
// it returns on the operands (for MUL the corner products of the smaller box lie between the
// larger box's extreme corner products; min, max, sqrt, floor, negation and the clamp of fix
// are monotone).  Evaluating the tape once on the host, with the device's operations, over a
// domain that contains every cell box and t-window therefore bounds every value any cell can
// compute.  Where a record's value (before fix) stays inside [-1e6, 1e6] on the domain it is
// never NaN or out of range in a cell, and fix leaves it unchanged.  sin / cos get the bound
// [-1.001, 1.001] (their deterministic routines stay within an ulp of [-1, 1]); every other form,
// and everything computed from it, gets none.  Bounds include 0: a record of a region that the
// pruned code skipped reads [+0, +0].
namespace {

constexpr float kBig = 1000000.0f;   // fix's bound (compute.wgsl:2534)

struct HBound { float lo, hi; bool ok; };

float hmin(float a, float b) { return __builtin_fminimum_numf(a, b); }
float hmax(float a, float b) { return __builtin_fmaximum_numf(a, b); }

HBound hb(float lo, float hi) {
    const bool ok = lo == lo && hi == hi && lo >= -kBig && hi <= kBig;
    return HBound{lo, hi, ok};
}

// the device rule of `form` (tsi_rules.h fwd_form_t, tsi_interval.h), before fix
HBound bound_rule(uint32_t form, HBound l, HBound r, float c) {
    const HBound none{0.0f, 0.0f, false};
    auto mul = [](float alo, float ahi, float blo, float bhi) {
        const float x = alo * blo, y = alo * bhi, z = ahi * blo, w = ahi * bhi;
        return hb(hmin(hmin(x, y), hmin(z, w)), hmax(hmax(x, y), hmax(z, w)));
    };
    switch (form) {
    case F_ADD_RR: return l.ok && r.ok ? hb(l.lo + r.lo, l.hi + r.hi) : none;
    case F_ADD_RL: return l.ok ? hb(l.lo + c, l.hi + c) : none;
    case F_SUB_RR: return l.ok && r.ok ? hb(l.lo - r.hi, l.hi - r.lo) : none;
    case F_SUB_RL: return l.ok ? hb(l.lo - c, l.hi - c) : none;
    case F_SUB_LR: return r.ok ? hb(c - r.hi, c - r.lo) : none;
    case F_MUL_RR: return l.ok && r.ok ? mul(l.lo, l.hi, r.lo, r.hi) : none;
    case F_MUL_RL: {
        if (!l.ok) return none;
        const float x = l.lo * c, y = l.hi * c;
        return c >= 0.0f ? hb(x, y) : hb(y, x);
    }
    case F_DIV_RL: {
        if (!l.ok || !(c < 0.0f || c > 0.0f)) return none;
        const float x = l.lo / c, z = l.hi / c;
        return hb(hmin(x, z), hmax(x, z));
    }
    case F_SQR: {
        if (!l.ok) return none;
        const float x = l.lo * l.lo, y = l.hi * l.hi;
        return hb(l.lo <= 0.0f && 0.0f <= l.hi ? 0.0f : hmin(x, y), hmax(x, y));
    }
    case F_SQRT: return l.ok ? hb(std::sqrt(hmax(l.lo, 0.0f)), std::sqrt(hmax(l.hi, 0.0f))) : none;
    case F_MIN_RR: return l.ok && r.ok ? hb(hmin(l.lo, r.lo), hmin(l.hi, r.hi)) : none;
    case F_MIN_RL: return l.ok ? hb(hmin(l.lo, c), hmin(l.hi, c)) : none;
    case F_MAX_RR: return l.ok && r.ok ? hb(hmax(l.lo, r.lo), hmax(l.hi, r.hi)) : none;
    case F_MAX_RL: return l.ok ? hb(hmax(l.lo, c), hmax(l.hi, c)) : none;
    case F_NEG: return l.ok ? hb(-l.hi, -l.lo) : none;
    case F_ABS: {
        if (!l.ok) return none;
        if (l.lo <= 0.0f && 0.0f <= l.hi) return hb(0.0f, hmax(-l.lo, l.hi));
        return l.hi < 0.0f ? hb(-l.hi, -l.lo) : l;
    }
    case F_RET: return l;
    case F_FLOOR: return l.ok ? hb(std::floor(l.lo), std::floor(l.hi)) : none;
    case F_SIN: case F_COS: return l.ok ? hb(-1.001f, 1.001f) : none;
    case F_NOP_RR: return hb(c, c);
    case F_ZERO: return hb(0.0f, 0.0f);
    default: return none;
    }
}

}  // namespace

std::vector<char> fix_free(const tsi_gpu_instruction *tape, uint32_t n, const std::vector<FwdIns> &fwd,
                           const JitDomain &d) {
    std::vector<HBound> b(n);
    std::vector<char> free(n, 0);
    for (uint32_t k = 0; k < 3; ++k) b[k] = hb(hmin(d.box[k], 0.0f), hmax(d.box[3 + k], 0.0f));
    b[3] = hb(hmin(d.t0, 0.0f), hmax(d.t1, 0.0f));
    for (uint32_t j = 4; j < n; ++j) {
        const tsi_gpu_instruction &g = tape[j];
        const uint32_t side = side_of(g);
        const HBound zero{0.0f, 0.0f, true};
        const HBound l = (side == 0 || side == 1) ? b[larg_of(g)] : zero;
        const HBound r = (side == 0 || side == 2) ? b[rarg_of(g)] : zero;
        const HBound v = side == 3 ? hb(0.0f, 0.0f) : bound_rule(fwd[j - 4].form, l, r, g.literal);
        free[j] = v.ok ? 1 : 0;
        // after fix (the identity where free), with 0 (a skipped region's value)
        b[j] = v.ok ? hb(hmin(v.lo, 0.0f), hmax(v.hi, 0.0f)) : HBound{0.0f, 0.0f, false};
    }
    return free;
}

JitLayout jit_layout(const std::vector<FwdIns> &fwd, const Lowered &L, bool prune, uint32_t group) {
    JitLayout y;
    y.group = group ? group : region_records(prune);
    y.regions = (L.nfwd + y.group - 1) / y.group;
    y.region_words = (y.regions + 31) / 32;
    for (uint32_t k = 0; k < L.nfwd; ++k) y.minmax += is_minmax_form(fwd[k].form) ? 1u : 0u;
    y.mode_words = (y.minmax + 15) / 16;
    y.record_words = 2 + y.region_words + y.mode_words;
    return y;
}

std::string jit_source(const tsi_gpu_instruction *tape, uint32_t n, const std::vector<FwdIns> &fwd,
                       const std::vector<RevIns> &rev, const Lowered &L, bool rev_kernel, bool prune,
                       const JitDomain *domain, bool skip, uint32_t group) {
    std::string o;
    o.reserve(size_t(n) * 500u);
    const JitLayout Y = jit_layout(fwd, L, prune, group);
    const uint32_t G = Y.group, RW = Y.region_words, MW = Y.mode_words;
    auto region_of = [&](uint32_t j) { return (j - 4) / G; };
    // MIN/MAX index of record j (records in order)
    std::vector<int> mmx(n, -1);
    for (uint32_t j = 4, m = 0; j < n; ++j)
        if (is_minmax_form(fwd[j - 4].form)) mmx[j] = int(m++);
    // records whose fix is the identity on the domain the kernels are specialised for
    const std::vector<char> ff = domain ? fix_free(tape, n, fwd, *domain) : std::vector<char>(n, 0);
    auto fixed = [&](uint32_t j, const std::string &e) { return ff[j] ? e : "fixk(" + e + ", K)"; };
    // Unpruned modules on a domain: records whose value has both bounds +0, positive or +inf,
    // never -0, negative or NaN (a square, an absolute value or a square root of a NaN-free
    // operand, sums of such values and of non-negative literals; fix keeps the property).  A
    // square root of such a value needs no clamp at 0 (SQRT_NN: the same bits as SQRT_I).
    std::vector<char> pz(n, 0);
    if (!prune && domain)
        for (uint32_t j = 4; j < n; ++j) {
            const uint32_t side = side_of(tape[j]), form = fwd[j - 4].form;
            const uint32_t a = larg_of(tape[j]), b = rarg_of(tape[j]);
            auto nan_free = [&](uint32_t k) { return k < 4u || ff[k]; };   // inputs: box coordinates, times
            const float c = tape[j].literal;
            if (form == F_SQR || form == F_ABS) pz[j] = side == 1 && nan_free(a);
            else if (form == F_SQRT) pz[j] = side == 1 && pz[a];
            else if (form == F_ADD_RR) pz[j] = side == 0 && pz[a] && pz[b];
            else if (form == F_ADD_RL) pz[j] = side == 1 && pz[a] && c >= 0.0f && !std::signbit(c) && c <= kBig;
        }
    // the forward expression of record j over operand expressions l, r
    auto fwd_expr = [&](uint32_t j, const std::string &l, const std::string &r) {
        const uint32_t form = fwd[j - 4].form;
        if (form == F_SQRT && side_of(tape[j]) == 1 && pz[larg_of(tape[j])]) return fixed(j, "SQRT_NN(" + l + ")");
        return fixed(j, "fwd_form_t<true>(" + std::to_string(form) + "u, " + l + ", " + r + ", " + lit(tape[j].literal) + ")");
    };
    // the variant being emitted: pruned (region mask, parent modes, choices) or full
    bool pr = prune;
    auto bit = [&](uint32_t r) {   // uniform test of region r's bit in the wave's region mask
        return pr ? "((rm" + std::to_string(r / 32) + " >> " + std::to_string(r % 32) + ") & 1u)"
                     : std::string("go");
    };
    auto open_region = [&](uint32_t r) {
        if (pr) return "    if (" + bit(r) + ") {\n";
        return std::string("    {\n        uint32_t go;\n        asm volatile(\"s_mov_b32 %0, 1\" : \"=s\"(go));\n"
                           "        if (go) {\n");
    };
    auto close_region = [&]() { return pr ? std::string("    }\n") : std::string("        }\n    }\n"); };
    auto mode_of = [&](uint32_t j) {
        const int m = mmx[j];
        return "((mw" + std::to_string(m / 16) + " >> " + std::to_string(2 * (m % 16)) + "u) & 3u)";
    };

    o += "// generated by tsi_jit.cpp: straight-line evaluation of one tape\n";
    o += "#include \"tsi_eval.h\"\n#include \"tsi_rules.h\"\n\nnamespace tsi {\n";
    o += "static constexpr uint32_t kTapeRecords = " + std::to_string(L.nfwd) + "u;\n";
    o += "static constexpr uint32_t kRegionWords = " + std::to_string(RW) + "u, kModeWords = " + std::to_string(MW) +
         "u;\n";
    o += "struct RegionMask { uint32_t w[" + std::to_string(RW) + "]; };\n";
    // fix() with its bounds in scalar registers: a 4-byte encoding per use instead of an
    // 8-byte one with a literal (the tape functions' code is streamed from L2 by every wave)
    o += "struct FixK { float lo, hi; };\n"
         "TSI_DI FixK fix_consts() {\n    uint32_t lo, hi;\n"
         "    asm volatile(\"s_mov_b32 %0, 0xc9742400\" : \"=s\"(lo));\n"
         "    asm volatile(\"s_mov_b32 %0, 0x49742400\" : \"=s\"(hi));\n"
         "    return FixK{__uint_as_float(lo), __uint_as_float(hi)};\n}\n"
         "TSI_DI Iv fixk(Iv x, FixK k) {   // constant operands keep fix() so that it folds\n"
         "    if (__builtin_constant_p(x.lo) && __builtin_constant_p(x.hi)) return fix(x);\n"
         "    return I(wmax(x.lo, k.lo), wmin(x.hi, k.hi));\n}\n";
    o += "TSI_DI uint32_t abs_code(Iv x) { return I_CONTAINS_ZERO(x) ? 3u : (x.hi < 0.0f ? 2u : 1u); }\n";
    o += "struct FwdOut { Iv field; uint32_t cost, records; };\n";
    // Unpruned modules (lean): the saved-operand stack as global memory, so that global loads are
    // counted one by one (vmcnt) and the next region's loads stay in flight while this region
    // waits for its own (through a generic pointer they are flat loads, after which only a wait
    // for all is possible), and each region's loads issued before the region above runs.
    // Pruned modules keep loads through the generic pointer, issued inside the region blocks.
    const bool lean = !prune;
    if (lean) o += "typedef __attribute__((address_space(1))) float2 gfloat2;\n";
    o += "struct RevOut { Iv s0, s1, s2, s3; uint32_t ops; };\n\n";

    // common prologue of the pruned functions: uniform region words, this lane's parent modes
    auto prologue = [&](bool need_modes) {
        std::string p;
        if (!pr) return p;
        for (uint32_t w = 0; w < RW; ++w)
            p += "    const uint32_t rm" + std::to_string(w) + " = __builtin_amdgcn_readfirstlane(mask.w[" +
                 std::to_string(w) + "]);\n";
        if (need_modes)
            for (uint32_t w = 0; w < MW; ++w)
                p += "    const uint32_t mw" + std::to_string(w) + " = prec ? prec[size_t(" + std::to_string(2 + RW + w) +
                     "u) * cap] : 0xFFFFFFFFu;\n";
        return p;
    };

    // ---- forward (compute.wgsl:3633-3772) ----
    // The tape functions are calls, not inlined into the level kernel: inlined, the register
    // allocation of the kernel's own state and of the tape interfere (several times the
    // registers of either alone, with spills); a call costs a few instructions per wavefront.
    // ---- the saved-operand stack of the reverse kernel ----
    // The reverse rules read operand values (MUL_RR, DIV, SQR, SQRT, the transcendental forms) or
    // only a decision (MIN/MAX: which operands receive [1, 1]; ABS: the sign class).  Each
    // value needed is saved once, when it is computed (row slot[k]); decisions are 2-bit codes,
    // 16 to a word, a word saved when its last code is set (row nslots + w).
    // A value of a one-register form with a literal (x + c, x - c, c - x, x * c, -x) whose
    // register operand is saved anyway, or shared by several such values, is recomputed in the
    // reverse pass from that operand (remat[k]: the operand) instead of being saved: the same
    // binary32 operations give the same bits, and a coordinate shared by many primitives is
    // then saved once.
    std::vector<int> slot(n, -1), cidx(n, -1), remat(n, -1);
    std::vector<char> recomp(n, 0);   // recomputed from its operands (unpruned modules)
    uint32_t nslots = 0, ncodes = 0;
    // SQRT records whose reverse rule takes the record's own value (sqrt_rr_of_value) instead of
    // SQRT_I of the saved operand: where fix is the identity, the saved value is that square
    // root, bit for bit (unpruned modules)
    std::vector<char> sqv(n, 0);
    for (uint32_t j = 4; j < n; ++j)
        sqv[j] = lean && ff[j] && fwd[j - 4].form == F_SQRT && side_of(tape[j]) == 1;
    auto remat_form = [&](uint32_t k) {
        const uint32_t f = fwd[k - 4].form;
        return k >= 4 && (f == F_ADD_RL || f == F_SUB_RL || f == F_MUL_RL || f == F_NEG || f == F_SUB_LR);
    };
    auto reg_operand = [&](uint32_t k) { return side_of(tape[k]) == 2 ? rarg_of(tape[k]) : larg_of(tape[k]); };
    {
        std::vector<char> need(n, 0);
        for (uint32_t j = 4; j < n; ++j) {
            const uint32_t side = side_of(tape[j]);
            if (side == 3) continue;
            const uint32_t form = fwd[j - 4].form;
            if (is_code_form(form)) { cidx[j] = int(ncodes++); continue; }
            bool nx = false, ny = false;
            rev_needs(form, nx, ny);
            if (sqv[j]) { need[j] = 1; continue; }
            if (nx) need[side == 2 ? rarg_of(tape[j]) : larg_of(tape[j])] = 1;
            if (ny) need[rarg_of(tape[j])] = 1;
        }
        if (lean && remat_budget() > 0) {
            // Unpruned modules: a needed value whose forward code, down to values that are saved
            // (the inputs, values of 4 or more users such as a coordinate shared by many
            // primitives, and values too costly to recompute), costs at most remat_budget() is
            // recomputed by the reverse pass with the same forward expression (the same binary32
            // operations, the same bits) instead of saved: the saved-operand stack is written by
            // every cell and read back by every cell, while the recomputation runs only where the
            // record's rule runs (not in a skipped subtree).
            std::vector<uint32_t> nusers(n, 0), rc(n, 0);
            std::vector<char> anchor(n, 1);
            for (uint32_t j = 4; j < n; ++j) {
                const uint32_t side = side_of(tape[j]);
                if (side == 0 || side == 1) nusers[larg_of(tape[j])] += 1;
                if (side == 0 || side == 2) nusers[rarg_of(tape[j])] += 1;
            }
            for (uint32_t k = 4; k < n; ++k) {   // ascending: the operands' costs are known
                const uint32_t side = side_of(tape[k]), ic = recompute_cost(fwd[k - 4].form);
                if (side == 3 || ic == 0) continue;
                uint32_t c = ic + (ff[k] ? 0u : 4u);
                if (side == 0 || side == 1) c += anchor[larg_of(tape[k])] ? 0u : rc[larg_of(tape[k])];
                if (side == 0 || side == 2) c += anchor[rarg_of(tape[k])] ? 0u : rc[rarg_of(tape[k])];
                rc[k] = c;
                anchor[k] = nusers[k] >= 4u || c > remat_budget();
            }
            for (uint32_t k = n; k-- > 4;) {   // descending: a recomputed value's operands become needed
                if (!need[k] || anchor[k]) continue;
                recomp[k] = 1;
                need[k] = 0;
                const uint32_t side = side_of(tape[k]);
                if (side == 0 || side == 1) need[larg_of(tape[k])] = 1;
                if (side == 0 || side == 2) need[rarg_of(tape[k])] = 1;
            }
        } else {
            std::vector<uint32_t> sharing(n, 0);
            for (uint32_t k = 4; k < n; ++k)
                if (need[k] && remat_form(k)) sharing[reg_operand(k)] += 1;
            for (uint32_t k = 4; k < n; ++k) {   // ascending: an operand is final when reached
                if (!need[k] || !remat_form(k)) continue;
                // the operand (a < k) is decided already: saved, recomputed, or not needed so far
                const uint32_t a = reg_operand(k);
                const bool a_saved = need[a] && remat[a] < 0;
                if (remat[a] < 0 && (a_saved || sharing[a] >= 2)) {
                    remat[k] = int(a);
                    need[k] = 0;
                    need[a] = 1;
                }
            }
        }
        for (uint32_t k = 0; k < n; ++k)
            if (need[k]) slot[k] = int(nslots++);
    }
    const uint32_t code_words = (ncodes + 15u) / 16u;
    // bit position of code c in its word: pruned modules place code i of a word at bits 2i;
    // unpruned ones shift each word left by 2 as its codes are appended, so the word's last code
    // sits at bits 0-1
    auto code_shift = [&](uint32_t c) -> uint32_t {
        if (!lean) return 2u * (c % 16u);
        const uint32_t in_word = std::min(16u, ncodes - 16u * (c / 16u));
        return 2u * (in_word - 1u - c % 16u);
    };
    // Unpruned modules save the code words two to a row (word w in row nslots + w / 2, .x for
    // even w), a row when its second word (or the last word) is complete; pruned modules one to a
    // row (.x).
    const uint32_t code_rows = lean ? (code_words + 1u) / 2u : code_words;
    // the region after which code word w is complete (lean: the rows, after their last word)
    std::vector<std::vector<uint32_t>> flush_after(Y.regions + 1);
    for (uint32_t j = 4; j < n; ++j)
        if (cidx[j] >= 0 && (cidx[j] % 16 == 15 || uint32_t(cidx[j]) == ncodes - 1u)) {
            const uint32_t w = uint32_t(cidx[j]) / 16u;
            if (!lean) flush_after[region_of(j)].push_back(w);
            else if (w % 2u == 1u || w + 1u == code_words) flush_after[region_of(j)].push_back(w / 2u);
        }
    if (rev_kernel && nslots + code_rows > L.stack + 1u)
        throw std::runtime_error("saved-operand stack layout exceeds the allocation");
    auto code_expr = [&](uint32_t j, const std::string &l, const std::string &rhs) {
        const uint32_t form = fwd[j - 4].form;
        if (form == F_ABS) return "abs_code(" + l + ")";
        const bool is_min = form == F_MIN_RR || form == F_MIN_RL;
        return std::string(is_min ? "min_choice(" : "max_choice(") + l + ", " + rhs + ")";
    };

    // choices: the MIN/MAX choices go to chbuf (always with pr)
    auto emit_forward = [&](const std::string &fname, bool choices) {
    o += "template <bool PUSH>\n__device__ __noinline__ FwdOut " + fname + "(Iv v0, Iv v1, Iv v2, Iv v3, float2 *" +
         (lean ? "stk_" : "stk") + ", uint32_t lane, const uint32_t *prec, uint32_t cap, uint32_t *chbuf, RegionMask mask) {\n";
    if (lean) o += "    gfloat2 *stk = (gfloat2 *)stk_;\n";
    o += "    const Iv ZI = I_FROM(0.0f);\n    const FixK K = fix_consts();\n    uint32_t cost = 0, recs = 0;\n";
    o += prologue(true);
    if (choices)   // skipped regions leave their MIN/MAX undecided (3)
        for (uint32_t w = 0; w < MW; ++w) o += "    uint32_t cw" + std::to_string(w) + (pr ? " = 0xFFFFFFFFu;\n" : " = 0u;\n");
    for (uint32_t j = 4; j < n; ++j) o += "    Iv " + V(j) + (pr ? " = ZI;\n" : ";\n");
    for (uint32_t w = 0; w < code_words; ++w) o += "    uint32_t kw" + std::to_string(w) + " = 0u;\n";
    for (uint32_t k = 0; k < 4; ++k)
        if (slot[k] >= 0) o += "    if (PUSH) " + P(uint32_t(slot[k])) + " = make_float2(" + V(k) + ".lo, " + V(k) + ".hi);\n";
    for (uint32_t j = 4; j < n; ++j) {
        const auto &g = tape[j];
        const FwdIns &f = fwd[j - 4];
        const uint32_t side = side_of(g);
        const uint32_t r = region_of(j);
        if ((j - 4) % G == 0) {
            o += open_region(r);
            uint32_t cost = 0, recs = 0;
            for (uint32_t k = j; k < n && region_of(k) == r; ++k, ++recs) cost += fwd[k - 4].cost;
            o += "        cost += " + std::to_string(cost) + "u; recs += " + std::to_string(recs) + "u;\n";
        }
        const std::string l = (side == 0 || side == 1) ? V(larg_of(g)) : "ZI";
        const std::string rr = (side == 0 || side == 2) ? V(rarg_of(g)) : "ZI";
        const std::string c = lit(g.literal);
        const std::string form = std::to_string(f.form) + "u";
        if (pr && mmx[j] >= 0) {
            // the parent's decided choice makes the MIN/MAX a copy (tsi_interp.h fwd_step)
            const std::string J = std::to_string(j);
            const bool rl = f.form == F_MIN_RL || f.form == F_MAX_RL;
            const bool is_min = f.form == F_MIN_RR || f.form == F_MIN_RL;
            const std::string rhs = rl ? "I_FROM(" + c + ")" : rr;
            o += "        const uint32_t md" + J + " = " + mode_of(j) + ";\n";
            o += "        " + V(j) + " = " + fixed(j, "sel(md" + J + " == 1u, " + l + ", sel(md" + J + " == 2u, " + rhs +
                 ", fwd_form_t<true>(" + form + ", " + l + ", " + rr + ", " + c + ")))") + ";\n";
            const int m = mmx[j];
            const std::string w = std::to_string(m / 16), sh = std::to_string(2 * (m % 16)) + "u";
            o += "        const uint32_t ch" + J + " = md" + J + " != 3u ? md" + J + " : " +
                 (is_min ? "min_choice(" : "max_choice(") + l + ", " + rhs + ");\n";
            o += "        cw" + w + " = (cw" + w + " & ~(3u << " + sh + ")) | (ch" + J + " << " + sh + ");\n";
        } else if (choices && mmx[j] >= 0) {
            const bool rl = f.form == F_MIN_RL || f.form == F_MAX_RL;
            const bool is_min = f.form == F_MIN_RR || f.form == F_MIN_RL;
            const std::string rhs = rl ? "I_FROM(" + c + ")" : rr;
            o += "        " + V(j) + " = " + fixed(j, "fwd_form_t<true>(" + form + ", " + l + ", " + rr + ", " + c + ")") + ";\n";
            const int m = mmx[j];
            const std::string w = std::to_string(m / 16), sh = std::to_string(2 * (m % 16)) + "u";
            o += "        cw" + w + " |= " + (is_min ? "min_choice(" : "max_choice(") + l + ", " + rhs + ") << " + sh + ";\n";
        } else {
            o += "        " + V(j) + " = " + fwd_expr(j, l, rr) + ";\n";
        }
        if (slot[j] >= 0)
            o += "        if (PUSH) " + P(uint32_t(slot[j])) + " = make_float2(" + V(j) + ".lo, " + V(j) + ".hi);\n";
        if (cidx[j] >= 0) {
            const bool rl = f.form == F_MIN_RL || f.form == F_MAX_RL;
            const std::string rhs = rl ? "I_FROM(" + c + ")" : rr;
            // a decided MIN/MAX of the pruned variant keeps the parent's choice (its code is
            // not read: the reverse pass takes the mode)
            const std::string code = (pr && mmx[j] >= 0) ? "ch" + std::to_string(j) : code_expr(j, l, rhs);
            if (lean)   // appended: the constants stay inline (no shifted select constants to materialise)
                o += "        if (PUSH) kw" + std::to_string(cidx[j] / 16) + " = (kw" + std::to_string(cidx[j] / 16) +
                     " << 2u) | " + code + ";\n";
            else
                o += "        if (PUSH) kw" + std::to_string(cidx[j] / 16) + " |= " + code + " << " +
                     std::to_string(2 * (cidx[j] % 16)) + "u;\n";
        }
        if ((j - 4) % G == G - 1 || j == n - 1) {
            o += close_region();
            for (uint32_t w : flush_after[region_of(j)]) {
                if (!lean) {
                    o += "    if (PUSH) " + P(nslots + w) + " = make_float2(__uint_as_float(kw" + std::to_string(w) +
                         "), 0.0f);\n";
                    continue;
                }
                const std::string hi_word = 2u * w + 1u < code_words ? "__uint_as_float(kw" + std::to_string(2u * w + 1u) + ")"
                                                                      : std::string("0.0f");
                o += "    if (PUSH) " + P(nslots + w) + " = make_float2(__uint_as_float(kw" + std::to_string(2u * w) +
                     "), " + hi_word + ");\n";
            }
        }
    }
    if (choices)
        for (uint32_t w = 0; w < MW; ++w)
            o += "    chbuf[" + std::to_string(w * 64u) + "u] = cw" + std::to_string(w) + ";\n";
    o += "    return FwdOut{" + V(n - 1) + ", cost, recs};\n}\n\n";
    };

    // ---- reverse (compute.wgsl:3780-3964) ----
    // sens of instruction k is s<k>, [+0, +0] until its first contribution (j descending).
    // Temporaries of record j carry the suffix _j.  Decided MIN/MAX pass the contribution to
    // the kept operand only (tsi_interp.h rev_step).
    auto emit_reverse = [&](const std::string &fname) {
        o += "__device__ __noinline__ RevOut " + fname + "(const float2 *" + (lean ? "stk_" : "stk") +
             ", uint32_t lane, const uint32_t *prec, uint32_t cap, RegionMask mask) {\n";
        if (lean) o += "    const gfloat2 *stk = (const gfloat2 *)stk_;\n";
        o += "    const Iv ZI = I_FROM(0.0f), ONE = I_FROM(1.0f);\n    const FixK K = fix_consts();\n    uint32_t ops = 0;\n";
        o += prologue(true);
        std::vector<char> touched(n, 0), live(n, 0);
        for (uint32_t j = 4; j < n; ++j) {
            const auto &g = tape[j];
            const uint32_t side = side_of(g);
            if (side == 0 || side == 1) touched[larg_of(g)] = 1;
            if (side == 0 || side == 2) touched[rarg_of(g)] = 1;
        }
        // Short-circuit (compute.wgsl:3802-3805): a single-use operand of a record whose
        // sensitivity is [0, 0] gets [0, 0].  After fix, the derivative of these forms always
        // has a finite bound (it is ordered, or a decision code), so the product with a zero
        // sensitivity is already +-0: the select would only fix the sign of a zero, and it is
        // left out (the work count stays).
        auto zero_product = [&](uint32_t form) {
            return is_code_form(form) || form == F_ADD_RR || form == F_ADD_RL || form == F_SUB_RR ||
                   form == F_SUB_RL || form == F_SUB_LR || form == F_NEG || form == F_RET || form == F_MUL_RL ||
                   form == F_DIV_RL || form == F_MUL_RR || form == F_SQR || form == F_SQRT || form == F_DIV_RR ||
                   form == F_DIV_LR || form == F_MOD_RL;
        };
        auto single_use = [&](uint32_t j) {   // the rule's operands all single-use (short-circuit applies)
            const RevIns &rv = rev[(n - 1) - j];
            const uint32_t side = side_of(tape[j]);
            if (side == 1) return (rv.flags & R_U1_A) != 0;
            if (side == 2) return (rv.flags & R_U1_B) != 0;
            return (rv.flags & R_U1_A) && (rv.flags & R_U1_B);
        };
        // Skipped subtrees (unpruned modules).  Record g dominates record k when every path from
        // the output to k passes through g: k's sensitivity is then a sum of products along paths
        // through g's.  With g's sensitivity [0, 0] and every record that g dominates either
        // short-circuited or of a form whose zero-sensitivity product is +-0, every sensitivity
        // below g is +-0 and every contribution leaving the subtree is +-0: running the subtree
        // changes no value but the sign of a zero (on which no output depends) and counts work
        // only for the records that cannot be short-circuited (the region's static count, kept).
        // Such subtrees of at least skip_min_records() records are skipped by a uniform branch
        // when g's sensitivity is [0, 0] in every lane (lv<g>, set when g's rule runs); each
        // record's code is guarded by its nearest dominating guard only, since a guard below a
        // skipped one keeps lv = false.  Sensitivities of records outside the subtree (and of the
        // inputs, which the function returns) whose first contribution is skipped get [0, 0].
        std::vector<int> idom(n, -1), depth(n, 0), gsize(n, 1), guard(n, -1);
        std::vector<char> reach(n, 0), desc_unsafe(n, 0), is_guard(n, 0);
        if (lean && skip && skip_min_records() > 0 && n > 5) {
            std::vector<std::vector<uint32_t>> users(n);
            for (uint32_t c = 4; c < n; ++c) {
                const uint32_t side = side_of(tape[c]);
                if (side == 0 || side == 1) users[larg_of(tape[c])].push_back(c);
                if (side == 0 || side == 2) users[rarg_of(tape[c])].push_back(c);
            }
            reach[n - 1] = 1;
            for (uint32_t j = n - 1; j-- > 0;) {
                int d = -1;
                for (uint32_t c : users[j]) {
                    if (!reach[c]) continue;
                    int a = int(c);
                    if (d < 0) { d = a; continue; }
                    while (a != d) {   // nearest common dominator of the users
                        if (depth[a] >= depth[d]) a = idom[a];
                        else d = idom[d];
                    }
                }
                if (d < 0) continue;
                reach[j] = 1;
                idom[j] = d;
                depth[j] = depth[d] + 1;
            }
            for (uint32_t j = 0; j < n; ++j) {   // ascending: a record's subtree is complete when reached
                if (!reach[j] || idom[j] < 0) continue;
                const bool unsafe = j >= 4 && side_of(tape[j]) != 3 && !single_use(j) && !zero_product(fwd[j - 4].form);
                desc_unsafe[idom[j]] |= char(unsafe || desc_unsafe[j]);
                gsize[idom[j]] += gsize[j];
            }
            // guards: operands of a MIN/MAX, whose decided choice gives them [0, 0] while the
            // MIN/MAX itself has a sensitivity (elsewhere a zero below a nonzero sensitivity
            // needs an exact cancellation, and the test would only cost)
            for (uint32_t g = 4; g + 1 < n; ++g) {
                bool choice_operand = false;
                for (uint32_t c : users[g]) {
                    const uint32_t fc = fwd[c - 4].form;
                    choice_operand |= fc == F_MIN_RR || fc == F_MAX_RR || fc == F_MIN_RL || fc == F_MAX_RL;
                }
                is_guard[g] = reach[g] && !desc_unsafe[g] && choice_operand &&
                              uint32_t(gsize[g] - 1) >= skip_min_records();
            }
            for (uint32_t j = n - 1; j-- > 0;)   // descending: a record's dominators come first
                if (reach[j] && idom[j] >= 0) guard[j] = is_guard[idom[j]] ? idom[j] : guard[idom[j]];
            for (uint32_t g = 4; g < n; ++g)
                if (is_guard[g]) o += "    bool lv" + std::to_string(g) + " = false;\n";
        }
        auto dominated_by = [&](uint32_t k, uint32_t g) {
            int a = int(k);
            while (a >= 0 && depth[a] > depth[g]) a = idom[a];
            return a == int(g);
        };
        for (uint32_t k = 0; k < n; ++k)
            if (slot[k] >= 0) o += "    float2 sv" + std::to_string(k) + ";\n";
        for (uint32_t w = 0; w < code_words; ++w) o += "    uint32_t kr" + std::to_string(w) + ";\n";
        for (uint32_t k = 0; k < n; ++k)
            if (touched[k]) o += "    Iv " + S(k) + (pr ? " = ZI;\n" : ";\n");
        uint64_t static_ops = 0;
        // operand x / y of record j in the reverse pass: its saved value, or [+0, +0] when the
        // pruned forward pass skipped the value's region (as the value itself was then)
        auto operand_of = [&](uint32_t j, bool second) -> int {
            const uint32_t side = side_of(tape[j]);
            bool nx = false, ny = false;
            rev_needs(fwd[j - 4].form, nx, ny);
            if (sqv[j]) return second ? -1 : int(j);   // the record's own value
            if (!second) return nx ? int(side == 2 ? rarg_of(tape[j]) : larg_of(tape[j])) : -1;
            return ny ? int(rarg_of(tape[j])) : -1;
        };
        std::function<std::string(uint32_t, int)> saved = [&](uint32_t j, int k) -> std::string {
            if (k < 0) return "ZI";
            std::string v;
            if (recomp[k]) {   // recomputed: the forward expression over recomputed or saved operands
                const uint32_t side = side_of(tape[k]);
                const std::string l = (side == 0 || side == 1) ? saved(j, int(larg_of(tape[k]))) : "ZI";
                const std::string r = (side == 0 || side == 2) ? saved(j, int(rarg_of(tape[k]))) : "ZI";
                v = fwd_expr(uint32_t(k), l, r);
            } else if (remat[k] >= 0) {   // recomputed from its saved register operand
                const std::string x = saved(j, remat[k]);
                const bool lr = side_of(tape[k]) == 2;
                v = fixed(uint32_t(k), "fwd_form_t<true>(" + std::to_string(fwd[k - 4].form) + "u, " +
                                           (lr ? "ZI, " + x : x + ", ZI") + ", " + lit(tape[k].literal) + ")");
            } else {
                v = "I(sv" + std::to_string(k) + ".x, sv" + std::to_string(k) + ".y)";
            }
            if (pr && k >= 4 && region_of(uint32_t(k)) != region_of(j))
                return "(" + bit(region_of(uint32_t(k))) + " ? " + v + " : ZI)";
            return v;
        };
        uint32_t nrow_loads = 0;   // loads of code-word rows (each its own temporary)
        // the saved values a value is read or recomputed from
        std::function<void(uint32_t, std::vector<char> &)> leaves = [&](uint32_t k, std::vector<char> &out) {
            if (!recomp[k]) { out[k] = 1; return; }
            const uint32_t side = side_of(tape[k]);
            if (side == 0 || side == 1) leaves(larg_of(tape[k]), out);
            if (side == 0 || side == 2) leaves(rarg_of(tape[k]), out);
        };
        // saved values (sv) and code words (kr) read by records hi-1 .. lo
        auto region_needs = [&](uint32_t hi, uint32_t lo, std::vector<char> &vneed, std::vector<char> &wneed) {
            vneed.assign(n, 0);
            wneed.assign(code_words + 1, 0);
            for (uint32_t jj = hi; jj-- > lo;) {
                if (side_of(tape[jj]) == 3) continue;
                if (cidx[jj] >= 0) {
                    wneed[uint32_t(cidx[jj]) / 16u] = 1;
                    continue;
                }
                for (bool second : {false, true}) {
                    int k = operand_of(jj, second);
                    if (k >= 0 && remat[k] >= 0) k = remat[k];
                    if (k >= 0) leaves(uint32_t(k), vneed);
                }
            }
        };
        // loads of what records hi-1 .. lo read, except what the region running now reads too
        // (skip): that is in registers already (the stack is not written during the reverse
        // pass), and reloading it here would make this region wait for the new load
        auto emit_loads = [&](uint32_t hi, uint32_t lo, const std::vector<char> *vskip = nullptr,
                              const std::vector<char> *wskip = nullptr) {
            if (!lean) {   // pruned modules: in record order, each value once per call
                std::vector<char> vdone(n, 0), wdone(code_words + 1, 0);
                for (uint32_t jj = hi; jj-- > lo;) {
                    if (side_of(tape[jj]) == 3) continue;
                    if (cidx[jj] >= 0) {
                        const uint32_t w = uint32_t(cidx[jj]) / 16u;
                        if (!wdone[w])
                            o += "        kr" + std::to_string(w) + " = __float_as_uint(" + P(nslots + w) + ".x);\n";
                        wdone[w] = 1;
                        continue;
                    }
                    for (bool second : {false, true}) {
                        int k = operand_of(jj, second);
                        if (k >= 0 && remat[k] >= 0) k = remat[k];
                        if (k >= 0 && !vdone[k]) o += "        sv" + std::to_string(k) + " = " + P(uint32_t(slot[k])) + ";\n";
                        if (k >= 0) vdone[k] = 1;
                    }
                }
                return;
            }
            std::vector<char> vneed, wneed;
            region_needs(hi, lo, vneed, wneed);
            // a row's two words with one load; a word the regions running before its use read
            // too (skip) is in a register already and is not assigned again
            for (uint32_t m = 0; 2u * m < code_words; ++m) {
                const bool a = wneed[2u * m] && !(wskip && (*wskip)[2u * m]);
                const bool b = 2u * m + 1u < code_words && wneed[2u * m + 1u] && !(wskip && (*wskip)[2u * m + 1u]);
                if (!a && !b) continue;
                const std::string t = "crow" + std::to_string(nrow_loads++);
                o += "        const float2 " + t + " = " + P(nslots + m) + ";\n";
                if (a) o += "        kr" + std::to_string(2u * m) + " = __float_as_uint(" + t + ".x);\n";
                if (b) o += "        kr" + std::to_string(2u * m + 1u) + " = __float_as_uint(" + t + ".y);\n";
            }
            for (uint32_t k = n; k-- > 0;)
                if (vneed[k] && !(vskip && (*vskip)[k]))
                    o += "        sv" + std::to_string(k) + " = " + P(uint32_t(slot[k])) + ";\n";
        };
        // regions from the top: records [hi, lo) with lo the region start
        auto region_lo = [&](uint32_t hi) { return 4u + ((hi - 5u) / G) * G; };
        // The loads sit between the region blocks, outside their conditional bodies: assigned
        // inside one, a value would merge with its old value after the block, and the copy that
        // merge needs would wait for the load there.  Regions that a pruned program skips get
        // their operands loaded all the same (the rows exist; the values are then unused).
        // lean: the regions from the top, and the loads of the region D below the one about to
        // run, except what the D regions from this one on read (those are in registers already)
        std::vector<std::pair<uint32_t, uint32_t>> regs;   // (hi, lo)
        for (uint32_t hi = n; hi > 4u; hi = region_lo(hi)) regs.push_back({hi, region_lo(hi)});
        const uint32_t D = prefetch_regions();
        auto lean_loads = [&](size_t i) {   // loads of region i, skipping regions i-D .. i-1
            std::vector<char> vskip(n, 0), wskip(code_words + 1, 0), vn, wn;
            for (size_t q = i >= D ? i - D : 0; q < i; ++q) {
                region_needs(regs[q].first, regs[q].second, vn, wn);
                for (uint32_t k = 0; k < n; ++k) vskip[k] |= vn[k];
                for (uint32_t w = 0; w <= code_words; ++w) wskip[w] |= wn[w];
            }
            emit_loads(regs[i].first, regs[i].second, &vskip, &wskip);
        };
        if (lean) {   // the top regions' operands
            for (size_t i = 0; i < D && i < regs.size(); ++i) lean_loads(i);
        } else if (!pr) {
            o += open_region(0);
            emit_loads(n, region_lo(n));
            o += close_region();
        }
        size_t ri = 0;
        for (uint32_t hi = n; hi > 4u; hi = region_lo(hi), ++ri) {
            const uint32_t lo = region_lo(hi), r = region_of(lo);
            // the operands of the region D below go in flight before this region runs
            if (lean && ri + D < regs.size()) lean_loads(ri + D);
            o += open_region(r);
            if (!lean) {
                if (pr) {
                    // operands of this region: loaded here unless the region above (run before
                    // it) prefetched them
                    if (hi == n) emit_loads(hi, lo);
                    else {
                        o += "        if (!" + bit(region_of(hi)) + ") {\n";
                        emit_loads(hi, lo);
                        o += "        }\n";
                    }
                }
                if (lo > 4u) emit_loads(lo, region_lo(lo));   // the next region's operands go in flight
            }
            // the open guarded block (its guard, the sensitivities first assigned in it)
            int cur = -1;
            std::vector<uint32_t> first;
            auto close_block = [&]() {
                if (cur < 0) return;
                std::string inits;
                for (uint32_t k : first)
                    if (k < 4u || !dominated_by(k, uint32_t(cur))) inits += "        " + S(k) + " = ZI;\n";
                o += inits.empty() ? std::string("        }\n") : "        } else {\n" + inits + "        }\n";
                cur = -1;
                first.clear();
            };
            auto assigned = [&](uint32_t k) {
                if (cur >= 0 && !live[k]) first.push_back(k);
                live[k] = 1;
            };
            for (uint32_t j = hi - 1; j >= lo; --j) {
                const auto &g = tape[j];
                const FwdIns &f = fwd[j - 4];
                const RevIns &rv = rev[(n - 1) - j];
                const uint32_t side = side_of(g);
                if (side != 3) {
                    if (guard[j] != cur) {
                        close_block();
                        if (guard[j] >= 0) {
                            cur = guard[j];
                            o += "        if (lv" + std::to_string(cur) + ") {\n";
                        }
                    }
                    if (is_guard[j])
                        o += "        lv" + std::to_string(j) + " = " +
                             (live[j] ? "__builtin_amdgcn_ballot_w64(!(" + S(j) + ".lo == 0.0f && " + S(j) +
                                            ".hi == 0.0f)) != 0"
                                      : std::string("false")) + ";\n";
                    const uint32_t a = larg_of(g), b = rarg_of(g);
                    const std::string J = "_" + std::to_string(j);
                    std::string sj, sj0;
                    if (j == n - 1) { sj = "ONE"; sj0 = "false"; }
                    else if (live[j]) { sj = S(j); sj0 = "(" + S(j) + ".lo == 0.0f && " + S(j) + ".hi == 0.0f)"; }
                    else { sj = "ZI"; sj0 = "true"; }
                    const std::string x = saved(j, operand_of(j, false)), y = saved(j, operand_of(j, true));
                    const std::string form = std::to_string(f.form) + "u";
                    const bool mm = pr && mmx[j] >= 0;
                    // the rule's derivatives: from the saved values, or from the decision code
                    std::string rule = "rev_form_t<true>(" + form + ", " + x + ", " + y + ", " + lit(g.literal) + ")";
                    if (sqv[j]) rule = "Iv2{sqrt_rr_of_value(" + x + "), ZI}";
                    if (cidx[j] >= 0) {
                        const uint32_t w = uint32_t(cidx[j]) / 16u;
                        o += "        const uint32_t kc" + J + " = (kr" + std::to_string(w) + " >> " +
                             std::to_string(code_shift(uint32_t(cidx[j]))) + "u) & 3u;\n";
                        const std::string kc = "kc" + J;
                        if (f.form == F_ABS)
                            rule = "Iv2{sel(" + kc + " == 3u, I(-1.0f, 1.0f), sel(" + kc + " == 2u, I_FROM(-1.0f), ONE)), ZI}";
                        else if (f.form == F_MIN_RL || f.form == F_MAX_RL)
                            rule = "Iv2{sel((" + kc + " & 1u) != 0u, ONE, ZI), ZI}";
                        else
                            rule = "Iv2{sel((" + kc + " & 1u) != 0u, ONE, ZI), sel((" + kc + " & 2u) != 0u, ONE, ZI)}";
                    }
                    if (mm) {
                        o += "        const uint32_t md" + J + " = " + mode_of(j) + ";\n";
                        o += "        Iv2 d" + J + " = " + rule + ";\n";
                        // decided: the kept operand receives [1, 1] (rev_step)
                        o += "        d" + J + ".a = sel(md" + J + " != 3u, ONE, d" + J + ".a);\n";
                        o += "        d" + J + ".b = sel(md" + J + " != 3u, ONE, d" + J + ".b);\n";
                    } else {
                        o += "        const Iv2 d" + J + " = " + rule + ";\n";
                    }
                    const std::string cost = std::to_string(unsigned(rv.cost)) + "u";
                    // MUL_RR's derivatives are the saved operand values; a record's value
                    // already went through fix, which is idempotent, so it is not applied again
                    // (arguments are not fixed in the forward pass and keep it)
                    auto fixed_d = [&](const char *dx) {
                        const std::string d = "d" + J + "." + dx;
                        if (f.form == F_MUL_RR && (dx[0] == 'a' ? b : a) >= 4u) return d;
                        if (cidx[j] >= 0) return d;   // 0, 1 or -1: fix leaves them
                        // on a domain: MUL_RR's derivatives are argument values inside it, and
                        // SQR's is 2x with |x| <= 1000 where x^2 stays inside [-1e6, 1e6]
                        if (domain && (f.form == F_MUL_RR || (f.form == F_SQR && ff[j]))) return d;
                        return "fixk(" + d + ", K)";
                    };
                    // the first contribution to a sensitivity is not added to [+0, +0]: that
                    // addition only turns a -0 into +0, and no output depends on the sign of a
                    // zero sensitivity (dx/dy/dz are stored with +0, the time error takes |df|)
                    auto contrib = [&](uint32_t k, const char *dx) {
                        const std::string prod = "MUL_II(" + fixed_d(dx) + ", " + sj + ")";
                        return live[k] ? "ADD_II(" + prod + ", " + S(k) + ")" : prod;
                    };
                    auto old = [&](uint32_t k) { return live[k] ? S(k) : std::string("ZI"); };
                    const bool zp = zero_product(f.form);
                    if (side == 1 || side == 2) {
                        const uint32_t k = side == 1 ? a : b;
                        const bool u1 = side == 1 ? (rv.flags & R_U1_A) != 0 : (rv.flags & R_U1_B) != 0;
                        std::string val = contrib(k, "a");
                        if (mm) val = "((md" + J + " & 1u) ? " + val + " : " + old(k) + ")";
                        if (u1) {
                            o += "        const bool z" + J + " = " + sj0 + ";\n";
                            o += "        ops += z" + J + " ? 0u : " + cost + ";\n";
                            if (!zp) val = "sel(z" + J + ", ZI, " + val + ")";
                        } else {
                            static_ops += rv.cost;
                        }
                        o += "        " + S(k) + " = " + val + ";\n";
                        assigned(k);
                    } else {   // side 0: both previous values are read before either is written
                        const bool u1 = (rv.flags & R_U1_A) && (rv.flags & R_U1_B);
                        std::string va = contrib(a, "a"), vb = contrib(b, "b");
                        if (mm) {
                            va = "((md" + J + " & 1u) ? " + va + " : " + old(a) + ")";
                            vb = "((md" + J + " >> 1) ? " + vb + " : " + old(b) + ")";
                        }
                        if (u1) {
                            o += "        const bool z" + J + " = " + sj0 + ";\n";
                            o += "        ops += z" + J + " ? 0u : " + cost + ";\n";
                            if (!zp) {
                                va = "sel(z" + J + ", ZI, " + va + ")";
                                vb = "sel(z" + J + ", ZI, " + vb + ")";
                            }
                        } else {
                            static_ops += rv.cost;
                        }
                        o += "        const Iv na" + J + " = " + va + ";\n";
                        o += "        const Iv nb" + J + " = " + vb + ";\n";
                        o += "        " + S(a) + " = na" + J + ";\n        " + S(b) + " = nb" + J + ";\n";
                        assigned(a);
                        assigned(b);
                    }
                }
                if (j == lo) break;
            }
            close_block();
            // work of the records that cannot be short-circuited, counted when the region runs
            if (static_ops) o += "        ops += " + std::to_string(static_ops) + "u;\n";
            static_ops = 0;
            o += close_region();
        }
        o += "    return RevOut{";
        for (uint32_t k = 0; k < 4; ++k) o += (live[k] ? S(k) : std::string("ZI")) + ", ";
        o += "ops};\n}\n\n";
    };

    // pruned modules carry three forward variants: pruned (the parents' records apply and skip
    // enough regions), full with the MIN/MAX choices recorded (some cell may write a liveness
    // record its children can use) and full; and two reverse ones, pruned and full
    emit_forward("tape_forward", prune);
    if (rev_kernel) emit_reverse("tape_reverse");
    if (prune) {
        pr = false;
        emit_forward("tape_forward_full", false);
        emit_forward("tape_forward_choices", true);
        if (rev_kernel) emit_reverse("tape_reverse_full");
        pr = true;
    }

    // ---- liveness: the records reachable from the output through this cell's own
    // MIN/MAX choices (compute.wgsl:2884-2903), as region bits and per-MIN/MAX modes for the
    // children's evaluation.  Regions skipped for the parent cannot hold live records.
    if (prune) {
        o += "__device__ __noinline__ void tape_liveness(const uint32_t *chbuf, uint32_t *rec, uint32_t cap, "
             "RegionMask mask) {\n";
        o += prologue(false);
        for (uint32_t w = 0; w < MW; ++w) {
            o += "    const uint32_t cw" + std::to_string(w) + " = chbuf[" + std::to_string(w * 64u) + "u];\n";
            o += "    uint32_t ow" + std::to_string(w) + " = 0xFFFFFFFFu;\n";
        }
        for (uint32_t k = 0; k < n; ++k) o += "    bool l" + std::to_string(k) + " = " + (k == n - 1 ? "true" : "false") + ";\n";
        for (uint32_t w = 0; w < RW; ++w) o += "    uint32_t rw" + std::to_string(w) + " = 0u;\n";
        auto region_lo = [&](uint32_t hi) { return 4u + ((hi - 5u) / G) * G; };
        for (uint32_t hi = n; hi > 4u; hi = region_lo(hi)) {
            const uint32_t lo = region_lo(hi), r = region_of(lo);
            o += open_region(r);
            const std::string gr = "g" + std::to_string(r);
            o += "        bool " + gr + " = false;\n";
            for (uint32_t j = hi - 1; j >= lo; --j) {
                const auto &g = tape[j];
                const uint32_t side = side_of(g);
                const std::string lj = "l" + std::to_string(j);
                o += "        " + gr + " = " + gr + " || " + lj + ";\n";
                if (mmx[j] >= 0) {
                    const int m = mmx[j];
                    const std::string w = std::to_string(m / 16), sh = std::to_string(2 * (m % 16)) + "u";
                    const std::string ch = "((cw" + w + " >> " + sh + ") & 3u)";
                    o += "        ow" + w + " = (ow" + w + " & ~(3u << " + sh + ")) | ((" + lj + " ? " + ch + " : 3u) << " +
                         sh + ");\n";
                    if (side == 0 || side == 1) {
                        const std::string la = "l" + std::to_string(larg_of(g));
                        o += "        " + la + " = " + la + " || (" + lj + " && (" + ch + " & 1u));\n";
                    }
                    if (side == 0 || side == 2) {
                        const std::string lb = "l" + std::to_string(rarg_of(g));
                        o += "        " + lb + " = " + lb + " || (" + lj + " && (" + ch + " >> 1));\n";
                    }
                } else {
                    if (side == 0 || side == 1) {
                        const std::string la = "l" + std::to_string(larg_of(g));
                        o += "        " + la + " = " + la + " || " + lj + ";\n";
                    }
                    if (side == 0 || side == 2) {
                        const std::string lb = "l" + std::to_string(rarg_of(g));
                        o += "        " + lb + " = " + lb + " || " + lj + ";\n";
                    }
                }
                if (j == lo) break;
            }
            o += "        rw" + std::to_string(r / 32) + " |= " + gr + " ? " + std::to_string(1u << (r % 32)) + "u : 0u;\n";
            o += close_region();
        }
        o += "    if (rec) {\n";
        for (uint32_t w = 0; w < RW; ++w)
            o += "        rec[size_t(" + std::to_string(2 + w) + "u) * cap] = rw" + std::to_string(w) + ";\n";
        for (uint32_t w = 0; w < MW; ++w)
            o += "        rec[size_t(" + std::to_string(2 + RW + w) + "u) * cap] = ow" + std::to_string(w) + ";\n";
        o += "    }\n}\n\n";
    }

    // ---- the evaluation engine of eval_level (tsi_eval.h) ----
    o += std::string("struct TapeEngine {\n    static constexpr bool kPrune = ") + (prune ? "true" : "false") + ";\n";
    o += "    static constexpr uint32_t kRegions = " + std::to_string(Y.regions) + "u;\n";
    o += "    uint32_t lane, recs;\n    const uint32_t *prec;\n    uint32_t *chbuf;\n    RegionMask mask;\n"
         "    uint32_t variant;   // this wavefront's: 0 full, 1 full + choices, 2 pruned\n";
    o += "    template <class KA>\n    TSI_DI TapeEngine(const KA &a, uint32_t wave, uint32_t l)\n"
         "        : lane(l), recs(kTapeRecords), prec(nullptr), variant(0) {\n"
         "        for (uint32_t w = 0; w < kRegionWords; ++w) mask.w[w] = 0xFFFFFFFFu;\n"
         "        chbuf = a.choice ? a.choice + size_t(wave) * kModeWords * 64u + l : nullptr;\n    }\n";
    o += "    template <class KA>\n    TSI_DI void select(const KA &a, int, uint32_t, bool pruned_ok, const uint32_t *prec_in, uint32_t pidx,\n"
         "                       bool need_choices) {\n"
         "        prec = nullptr;\n"
         "        variant = 0;\n"
         "        for (uint32_t w = 0; w < kRegionWords; ++w) mask.w[w] = 0xFFFFFFFFu;\n";
    if (prune)
        // the pruned variant costs about a third more than the full one per region it runs:
        // it is taken when at most prune_live_pct % of the regions are live (default 75)
        o += "        uint32_t live = 0;\n"
             "        if (pruned_ok) {\n"
             "            prec = prec_in + pidx;\n"
             "            for (uint32_t w = 0; w < kRegionWords; ++w) {\n"
             "                uint32_t v = prec[size_t(2u + w) * a.prune_cap];\n"
             "                v |= uint32_t(__shfl_xor(int(v), 8, 64));\n"
             "                v |= uint32_t(__shfl_xor(int(v), 16, 64));\n"
             "                v |= uint32_t(__shfl_xor(int(v), 32, 64));\n"
             "                mask.w[w] = v;\n"
             "                live += uint32_t(__popc(__builtin_amdgcn_readfirstlane(v)));\n"
             "            }\n        }\n"
             "        variant = pruned_ok && live * 100u <= kRegions * a.prune_live_pct ? 2u : (need_choices ? 1u : 0u);\n"
             "        if (variant != 2u) {\n"
             "            prec = nullptr;\n"
             "            for (uint32_t w = 0; w < kRegionWords; ++w) mask.w[w] = 0xFFFFFFFFu;\n"
             "        }\n";
    o += "    }\n";
    const std::string fcall = "(x, y, z, t, stk, lane, prec, a.prune_cap, chbuf, mask)";
    o += "    template <bool PUSH, class KA>\n"
         "    TSI_DI uint32_t forward(const KA &a, Iv x, Iv y, Iv z, Iv t, float2 *stk, Iv &field, Slots &, Slots &) {\n";
    if (prune)
        o += "        const FwdOut r = variant == 2u ? tape_forward<PUSH>" + fcall + "\n"
             "                       : variant == 1u ? tape_forward_choices<PUSH>" + fcall + "\n"
             "                                       : tape_forward_full<PUSH>" + fcall + ";\n";
    else
        o += "        const FwdOut r = tape_forward<PUSH>" + fcall + ";\n";
    o += "        field = r.field;\n        recs = r.records;\n        return r.cost;\n    }\n";
    const std::string rcall = "(stk, lane, prec, a.prune_cap, mask)";
    if (rev_kernel) {
        o += "    template <class KA>\n    TSI_DI uint32_t reverse(const KA &a, const float2 *stk, Iv sens[4], Slots &, Slots &) {\n";
        if (prune)
            o += "        const RevOut r = variant == 2u ? tape_reverse" + rcall + " : tape_reverse_full" + rcall + ";\n";
        else o += "        const RevOut r = tape_reverse" + rcall + ";\n";
        o += "        sens[0] = r.s0; sens[1] = r.s1; sens[2] = r.s2; sens[3] = r.s3;\n        return r.ops;\n    }\n";
    } else {
        o += "    template <class KA>\n    TSI_DI uint32_t reverse(const KA &, const float2 *, Iv sens[4], Slots &, Slots &) {\n"
             "        for (int k = 0; k < 4; ++k) sens[k] = I_FROM(0.0f);\n        return 0u;\n    }\n";
    }
    o += "    TSI_DI uint32_t records() const { return recs; }\n";
    o += "    TSI_DI bool choices_recorded() const { return variant != 0u; }\n";
    if (prune)
        o += "    template <class KA>\n    TSI_DI void liveness(const KA &a, uint32_t *rec) { tape_liveness(chbuf, rec, a.prune_cap, mask); }\n";
    else
        o += "    template <class KA>\n    TSI_DI void liveness(const KA &, uint32_t *) {}\n";
    o += "};\n\n}  // namespace tsi\n\n";
    o += std::string("extern \"C\" __global__ void __launch_bounds__(256) ") +
         (rev_kernel ? "tsi_tape_eval_rev" : "tsi_tape_eval_fwd") +
         "(tsi::KArgs a, int level, float time) {\n    tsi::eval_level<" + (rev_kernel ? "true" : "false") +
         ", tsi::TapeEngine>(a, level, time);\n}\n";
    return o;
}

namespace {

// The definitions of a generated module: what it holds inside namespace tsi, without the
// includes and the kernel.
std::string module_body(const std::string &src) {
    const std::string open = "namespace tsi {\n", close = "}  // namespace tsi\n";
    const size_t a = src.find(open);
    const size_t b = src.rfind(close);
    if (a == std::string::npos || b == std::string::npos || b < a) throw std::runtime_error("module layout");
    return src.substr(a + open.size(), b - a - open.size());
}

}  // namespace

std::string jit_source_level(const tsi_gpu_instruction *tape, uint32_t n, const std::vector<FwdIns> &fwd,
                             const std::vector<RevIns> &rev, const Lowered &L, bool prune, const JitDomain *domain) {
    const std::string f = jit_source(tape, n, fwd, rev, L, false, prune, domain);
    const std::string r = jit_source(tape, n, fwd, rev, L, true, prune, domain);
    std::string o = "// generated by tsi_jit.cpp: straight-line evaluation of one tape, both lists of a level\n";
    o += "#include \"tsi_eval.h\"\n#include \"tsi_rules.h\"\n\nnamespace tsi {\nnamespace jf {   // forward-only list\n";
    o += module_body(f);
    o += "}  // namespace jf\nnamespace jr {   // reverse list\n";
    o += module_body(r);
    o += "}  // namespace jr\n}  // namespace tsi\n\n";
    // one launch per level: the forward-only list and the reverse list run side by side, each
    // on its share of the grid (level_split), instead of one after the other
    const std::string lb = level_waves() ? "256, " + std::to_string(level_waves()) : "256";
    o += "extern \"C\" __global__ void __launch_bounds__(" + lb + ") tsi_tape_eval_level(tsi::KArgs a, int level, float time) {\n";
    if (level_prio())
        o += "    // issue ahead of other kernels' wavefronts on the same SIMD (the previous frame's trace)\n"
             "    if (level >= 0) __builtin_amdgcn_s_setprio(" + std::to_string(level_prio()) + ");\n";
    o += "    const uint32_t nb = gridDim.x, b = blockIdx.x, w = threadIdx.x >> 6;\n"
         "    // level -1: the frame's speculative pass, on the whole grid (reverse list code)\n"
         "    const uint32_t nf = level < 0 ? 0u : __builtin_amdgcn_readfirstlane(tsi::level_split(a, level, nb));\n"
         "    const uint32_t bwave = b * 4u + w;\n"
         "    if (b < nf)\n"
         "        tsi::eval_level<false, tsi::jf::TapeEngine>(a, level, time, bwave, nf * 4u, bwave);\n"
         "    else\n"
         "        tsi::eval_level<true, tsi::jr::TapeEngine>(a, level, time, (b - nf) * 4u + w, (nb - nf) * 4u, bwave);\n"
         "}\n";
    return o;
}

std::string jit_source_batch(const std::vector<BatchTape> &tapes, const JitDomain *domain) {
    std::string o = "// generated by tsi_jit.cpp: the persistent batch kernel over " + std::to_string(tapes.size()) +
                    " tapes\n";
    o += "#include \"tsi_batch.h\"\n#include \"tsi_rules.h\"\n\nnamespace tsi {\n";
    for (size_t i = 0; i < tapes.size(); ++i) {
        const BatchTape &t = tapes[i];
        // without the skipped subtrees: the batched workload's wide swept windows leave most
        // MIN/MAX undecided, and its 256 requests ran 6 % slower with them; 32-record regions
        // (with 64: 611 against 643 frames/s batched, served p50 2.01 against 1.80 ms)
        const bool skip = batch_skip();
        const uint32_t group = batch_region_records();
        const std::string f = jit_source(t.tape, t.n, *t.fwd, *t.rev, *t.low, false, false, domain, skip, group);
        const std::string r = jit_source(t.tape, t.n, *t.fwd, *t.rev, *t.low, true, false, domain, skip, group);
        o += "namespace tb" + std::to_string(i) + " {\nnamespace jf {   // forward-only list\n" + module_body(f);
        o += "}  // namespace jf\nnamespace jr {   // reverse list\n" + module_body(r);
        o += "}  // namespace jr\n}  // namespace tb" + std::to_string(i) + "\n";
    }
    // the tape of a unit is the session's, uniform over the wavefront
    o += "struct BatchTapes {\n"
         "    template <bool REV, class E, class KA>\n"
         "    static TSI_DI void run(const KA &a, int level, float time, uint32_t g, uint32_t count, float2 *stk,\n"
         "                           LevelTally &tl) {\n"
         "        E eng(a, 0u, lane_id());\n"
         "        Slots Rlo, Rhi;\n"
         "        eval_group<REV, E>(a, level, time, g, count, stk, eng, Rlo, Rhi, tl, REV && level < 0);\n"
         "    }\n"
         "    template <bool REV, class KA>\n"
         "    static TSI_DI void group(uint32_t tape, const KA &a, int level, float time, uint32_t g, uint32_t count,\n"
         "                             float2 *stk, LevelTally &tl) {\n"
         "        switch (tape) {\n";
    for (size_t i = 0; i < tapes.size(); ++i) {
        const std::string ns = "tb" + std::to_string(i);
        o += "        case " + std::to_string(i) + "u:\n"
             "            if constexpr (REV) run<true, " + ns + "::jr::TapeEngine>(a, level, time, g, count, stk, tl);\n"
             "            else run<false, " + ns + "::jf::TapeEngine>(a, level, time, g, count, stk, tl);\n"
             "            break;\n";
    }
    o += "        default: break;\n        }\n    }\n};\n}  // namespace tsi\n\n";
    // one block per CU of up to 16 wavefronts (4 per SIMD)
    o += "extern \"C\" __global__ void __launch_bounds__(1024) tsi_batch_run(tsi::BatchArgs b) {\n"
         "    tsi::batch_run<tsi::BatchTapes>(b);\n}\n";
    return o;
}

namespace {

// the unsigned decimal at s[i..], or -1
long long read_uint(const std::string &s, size_t i) {
    long long v = -1;
    while (i < s.size() && s[i] >= '0' && s[i] <= '9') v = (v < 0 ? 0 : v * 10) + (s[i++] - '0');
    return v;
}

}  // namespace

std::string jit_check_bounds(const std::string &src, uint32_t stack_rows, uint32_t record_words,
                             uint32_t mode_words) {
    // stack rows: (stk + 64 r u)[lane]
    for (size_t i = src.find("(stk + "); i != std::string::npos; i = src.find("(stk + ", i + 1)) {
        const long long off = read_uint(src, i + 7);
        if (off < 0 || off % 64 != 0 || off / 64 >= stack_rows)
            return "stack row " + std::to_string(off / 64) + " of " + std::to_string(stack_rows) + " at offset " +
                   std::to_string(i);
    }
    // liveness-record words: prec[size_t(w u) * cap] / rec[size_t(w u) * cap]
    for (const char *pat : {"prec[size_t(", "rec[size_t("}) {
        const size_t len = std::strlen(pat);
        for (size_t i = src.find(pat); i != std::string::npos; i = src.find(pat, i + 1)) {
            const long long w = read_uint(src, i + len);
            if (w < 0) continue;   // a word index computed at run time (bounded by its loop)
            if (uint64_t(w) >= record_words)
                return "record word " + std::to_string(w) + " of " + std::to_string(record_words);
        }
    }
    // choice columns: chbuf[64 w u]
    for (size_t i = src.find("chbuf["); i != std::string::npos; i = src.find("chbuf[", i + 1)) {
        const long long off = read_uint(src, i + 6);
        if (off < 0 || off % 64 != 0 || uint64_t(off / 64) >= (mode_words ? mode_words : 1u))
            return "choice column " + std::to_string(off / 64) + " of " + std::to_string(mode_words);
    }
    return std::string();
}

}  // namespace tsi

// ---------------------------------------------------------------- compilation

namespace tsi {

namespace {

uint64_t fnv1a(const std::string &s, uint64_t h = 1469598103934665603ull) {
    for (unsigned char c : s) {
        h ^= c;
        h *= 1099511628211ull;
    }
    return h;
}

// Compiler flags of the per-tape code: those of the library's own kernels (no FMA
// contraction, packed FP32 disabled, uniform regions kept as scalar branches).
const char *const kOptions[] = {
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
};
// Added for unpruned modules (those that read the stack as global memory): machine sinking would
// move each region's prefetched loads down into the region that reads them, after the region
// above, and their latency would be exposed again.
const char *const kLeanOptions[] = {"-mllvm", "-disable-machine-sink"};

std::string cache_dir() {
    if (const char *d = std::getenv("TSI_JIT_CACHE")) return d;
    std::string base;
    if (const char *x = std::getenv("XDG_CACHE_HOME")) base = x;
    else if (const char *h = std::getenv("HOME")) base = std::string(h) + "/.cache";
    else base = "/tmp";
    return base + "/tsi-jit";
}

bool read_file(const std::string &path, std::vector<char> &out) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    bool ok = sz > 0;
    if (ok) {
        out.resize(size_t(sz));
        ok = std::fread(out.data(), 1, out.size(), f) == out.size();
    }
    std::fclose(f);
    return ok;
}

void mkdir_p(const std::string &dir) {
    for (size_t i = 1; i <= dir.size(); ++i)
        if (i == dir.size() || dir[i] == '/') ::mkdir(dir.substr(0, i).c_str(), 0755);
}

void write_file_atomic(const std::string &dir, const std::string &name, const std::vector<char> &data) {
    mkdir_p(dir);   // failures only lose caching
    const std::string tmp = dir + "/." + name + "." + std::to_string(::getpid());
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    const bool ok = std::fwrite(data.data(), 1, data.size(), f) == data.size();
    std::fclose(f);
    if (ok) std::rename(tmp.c_str(), (dir + "/" + name).c_str());
    else std::remove(tmp.c_str());
}

std::mutex g_mu;
std::map<uint64_t, std::vector<char>> g_mem;   // code objects compiled by this process

}  // namespace

int jit_compile(const std::string &src, bool prune, std::vector<char> &code, std::string &log, bool &cache_hit,
                double &ms) {
    const auto t0 = std::chrono::steady_clock::now();
    cache_hit = false;
    log.clear();
    std::vector<const char *> opts(std::begin(kOptions), std::end(kOptions));
    if (!prune) opts.insert(opts.end(), std::begin(kLeanOptions), std::end(kLeanOptions));
    std::string key_text = src;
    for (const char *opt : opts) key_text += std::string("\n") + opt;
    // the headers the source can include: the batch kernel's header only for batch modules
    const bool batch = src.find("#include \"tsi_batch.h\"") != std::string::npos;
    for (int i = 0; i < tsi_jit_num_headers; ++i)
        if (batch || std::strcmp(tsi_jit_header_names[i], "tsi_batch.h") != 0) key_text += tsi_jit_header_texts[i];
    int ver_major = 0, ver_minor = 0;
    hiprtcVersion(&ver_major, &ver_minor);
    key_text += "\nhiprtc " + std::to_string(ver_major) + "." + std::to_string(ver_minor);
    const uint64_t key = fnv1a(key_text);
    char name[40];
    std::snprintf(name, sizeof name, "%016llx.co", (unsigned long long)key);
    // $TSI_JIT_DUMP: a directory that receives each module's generated source as <key>.hip
    if (const char *dump = std::getenv("TSI_JIT_DUMP")) {
        char src_name[40];
        std::snprintf(src_name, sizeof src_name, "%016llx.hip", (unsigned long long)key);
        write_file_atomic(dump, src_name, std::vector<char>(src.begin(), src.end()));
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_mem.find(key);
        if (it != g_mem.end()) {
            code = it->second;
            cache_hit = true;
        }
    }
    const std::string dir = cache_dir();
    if (!cache_hit && read_file(dir + "/" + name, code)) {
        cache_hit = true;
        ::utime((dir + "/" + name).c_str(), nullptr);   // marks the entry as in use (see the build's pruning)
    }
    if (!cache_hit) {
        hiprtcProgram prog;
        if (hiprtcCreateProgram(&prog, src.c_str(), "tsi_tape.hip", tsi_jit_num_headers, tsi_jit_header_texts,
                                tsi_jit_header_names) != HIPRTC_SUCCESS) {
            log = "hiprtcCreateProgram failed";
            return TSI_E_JIT;
        }
        const hiprtcResult rc = hiprtcCompileProgram(prog, int(opts.size()), opts.data());
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        if (ls > 1) {
            log.resize(ls);
            hiprtcGetProgramLog(prog, &log[0]);
        }
        if (rc != HIPRTC_SUCCESS) {
            hiprtcDestroyProgram(&prog);
            return TSI_E_JIT;
        }
        size_t cs = 0;
        hiprtcGetCodeSize(prog, &cs);
        code.resize(cs);
        hiprtcGetCode(prog, code.data());
        hiprtcDestroyProgram(&prog);
        write_file_atomic(dir, name, code);
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_mem[key] = code;
    }
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return TSI_OK;
}

}  // namespace tsi
