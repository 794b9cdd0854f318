// tsi_tape.h -- opcode numbering and the lowered (device) tape format, shared by the host
// lowering pass and the HIP kernels.
#pragma once
#ifdef __HIPCC_RTC__
using __hip_internal::uint8_t;
using __hip_internal::uint16_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#else
#include <stdint.h>
#endif
#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define TSI_HD __host__ __device__
#else
#define TSI_HD
#endif

namespace tsi {

// src/ir.rs:8-43 opcode numbering (kept: it is part of the 32-byte tape format).
namespace Op {
enum : uint32_t {
    NOP = 0, ARGX = 1, ARGY = 2, ARGZ = 3, ARGW = 4, ADD = 5, SUB = 6, MUL = 7, DIV = 8, EXP = 9, LOG = 10,
    SQR = 11, SQRT = 12, POW = 13, SIN = 14, COS = 15, NEG = 16, MIN = 17, MAX = 18, ABS = 19, CUP = 20,
    CAP = 21, BR = 22, PHI = 23, TRUNC = 24, FLOOR = 25, CEIL = 26, SIGN = 27, RET = 28, MOD = 29, ATAN2 = 30
};
}

// Dense "form" = (opcode, operand side) pairs the forward switch of compute.wgsl:3642-3764
// distinguishes.  RR = both registers, RL = register/literal, LR = literal/register.
enum Form : uint8_t {
    F_ADD_RR = 0, F_SUB_RR, F_MUL_RR, F_DIV_RR, F_MIN_RR, F_MAX_RR, F_POW_RR, F_MOD_RR, F_ATAN2_RR, F_NOP_RR,
    F_ADD_RL, F_SUB_RL, F_MUL_RL, F_DIV_RL, F_MIN_RL, F_MAX_RL, F_POW_RL, F_MOD_RL, F_ATAN2_RL,
    F_SQR, F_SQRT, F_EXP, F_LOG, F_NEG, F_ABS, F_FLOOR, F_SIN, F_COS, F_RET,
    F_SUB_LR, F_DIV_LR,
    F_ZERO,        // lit/lit: value [0,0] (compute.wgsl:3765 `default: {}`)
    F_NINF_RR,     // an opcode the forward switch does not know: I_FROM(-inf), both operands registers
    F_NINF_RL,
    F_NINF_LR,
    F_COUNT
};

// Register-file size of the interpreter (values, then sensitivities): slots per lane.
constexpr int kSlots = 32;
constexpr uint8_t kNoSlot = 0xFF;

// Forward instruction, 16 bytes.  Operands a/b and destination d are register-file slots.
// push: stack position (in intervals) where the operand values the reverse rule needs are
// saved; npush_side bits 0-1 = how many (0, 1 = left, 2 = left and right, 3 = right only),
// bits 2-3 = operand side (0 reg/reg, 1 reg/lit, 2 lit/reg, 3 lit/lit).  index = position of
// the record in the full program (tape instruction - 4).  modes: the MIN/MAX choice of each
// of the 8 parents a wavefront evaluates (2 bits each; 0xFFFF in the full program).
struct FwdIns {
    uint8_t form, a, b, d;
    uint16_t push;
    uint8_t npush_side;
    uint8_t cost;       // algorithmic f32 ops of the forward rule
    float lit;
    uint16_t index;
    uint16_t modes;
};

// Reverse instruction, 16 bytes (executed for j = n-1 down to 4; record r of the full
// reverse program belongs to forward record nfwd-1-r).
enum : uint8_t {
    R_SJ_ONE = 1,     // sens(j) = [1,1]: the output (compute.wgsl:3777)
    R_SJ_ZERO = 2,    // sens(j) never accumulated: [0,0]
    R_FT_A = 4,       // first touch of the left operand's sens: previous value is [+0,+0]
    R_FT_B = 8,
    R_U1_A = 16,      // uses(left operand) == 1 (short-circuit test, compute.wgsl:3802)
    R_U1_B = 32
};
struct RevIns {
    uint8_t form, sj, sa, sb;
    uint8_t flags;
    uint8_t side;
    uint16_t pop;     // stack position of the saved operands (same as FwdIns::push)
    float lit;
    uint8_t cost;     // algorithmic f32 ops when not short-circuited
    uint8_t npush;    // saved operand intervals at `pop` (FwdIns npush of the same instruction)
    uint16_t modes;   // as FwdIns::modes
};
static_assert(sizeof(FwdIns) == 16, "16-byte forward records");
static_assert(sizeof(RevIns) == 16, "16-byte reverse records");

enum { kPruneOff = 0, kPruneApprox = 1, kPruneExact = 2 };

// Records a program of m records occupies, zero padding included: the interpreter fetches
// 16-record groups ahead of execution (4 batches of 4 past the start of its last pair).
inline TSI_HD uint32_t program_stride(uint32_t m) { return (m + 7u) / 8u * 8u + 16u; }

// Host-side result of lowering a tape.
struct Lowered {
    uint32_t n = 0;                 // instructions in the tape
    uint32_t nfwd = 0, nrev = 0;    // = n - 4 each
    uint32_t stack = 0;             // pushed intervals per lane
    uint32_t slots = 0;             // max live values / sens
    uint32_t arg_slot[4] = {0, 1, 2, 3};
    uint32_t out_slot = 0;          // slot of the field value (instruction n-1)
    uint64_t fwd_cost = 0;          // algorithmic f32 ops of one forward pass
    uint32_t nminmax = 0;           // MIN/MAX records (choice words per lane: ceil(/16))
    int prune_class = 0;            // kPrune*
};

}  // namespace tsi
