// tsi_lower.cpp -- lowers a 32-byte GPUInstruction tape into the interpreter's programs.
//
// The reference evaluates every cell on a private copy of the whole tape in global memory
// (compute.wgsl:3622-3964: one 131 KB Tape per invocation, values and sens at
// [instruction]).  Here each lane keeps only the live values in a small register file:
//   forward:  linear-scan slot allocation over value lifetimes [j, last use];
//   reverse:  linear-scan slot allocation over sensitivity lifetimes [last user, j] walked
//             downwards, arguments pinned to slots 0..3;
//   stack:    operand values a reverse rule needs (compute.wgsl:2806-2970) are saved by the
//             forward pass at static positions and read back by the reverse pass.
// The j-descending accumulation order of compute.wgsl:3786 is kept exactly.
#include "tsi.h"
#include "tsi_tape.h"

#include <algorithm>
#include <vector>

namespace tsi {

namespace {

inline uint32_t op_of(const tsi_gpu_instruction &g) { return (g.opcode_uses >> 16) & ~(3u << 14); }
inline uint32_t side_of(const tsi_gpu_instruction &g) { return (g.opcode_uses >> 16) >> 14; }
inline uint32_t uses_of(const tsi_gpu_instruction &g) { return g.opcode_uses & 0xFFFFu; }
inline uint32_t larg_of(const tsi_gpu_instruction &g) { return g.larg_rarg >> 16; }
inline uint32_t rarg_of(const tsi_gpu_instruction &g) { return g.larg_rarg & 0xFFFFu; }

// compute.wgsl:3642-3764: which (op, side) pairs have a case; the rest evaluate to -inf.
Form form_of(uint32_t op, uint32_t side) {
    switch (side) {
    case 0:
        switch (op) {
        case Op::ADD: return F_ADD_RR;
        case Op::SUB: return F_SUB_RR;
        case Op::MUL: return F_MUL_RR;
        case Op::DIV: return F_DIV_RR;
        case Op::MIN: return F_MIN_RR;
        case Op::MAX: return F_MAX_RR;
        case Op::POW: return F_POW_RR;
        case Op::MOD: return F_MOD_RR;
        case Op::ATAN2: return F_ATAN2_RR;
        case Op::NOP: return F_NOP_RR;
        default: return F_NINF_RR;
        }
    case 1:
        switch (op) {
        case Op::SQR: return F_SQR;
        case Op::SQRT: return F_SQRT;
        case Op::EXP: return F_EXP;
        case Op::LOG: return F_LOG;
        case Op::NEG: return F_NEG;
        case Op::ABS: return F_ABS;
        case Op::FLOOR: return F_FLOOR;
        case Op::SIN: return F_SIN;
        case Op::COS: return F_COS;
        case Op::ADD: return F_ADD_RL;
        case Op::SUB: return F_SUB_RL;
        case Op::DIV: return F_DIV_RL;
        case Op::MUL: return F_MUL_RL;
        case Op::MIN: return F_MIN_RL;
        case Op::MAX: return F_MAX_RL;
        case Op::POW: return F_POW_RL;
        case Op::RET: return F_RET;
        case Op::MOD: return F_MOD_RL;
        case Op::ATAN2: return F_ATAN2_RL;
        default: return F_NINF_RL;
        }
    case 2:
        switch (op) {
        case Op::SUB: return F_SUB_LR;
        case Op::DIV: return F_DIV_LR;
        default: return F_NINF_LR;
        }
    default:
        return F_ZERO;
    }
}

// Operand intervals the reverse rule of a form reads: 0 none, 1 left, 2 both, 3 right.
uint8_t push_kind(Form f) {
    switch (f) {
    case F_SQR: case F_SQRT: case F_SIN: case F_COS: case F_EXP: case F_LOG: case F_FLOOR: case F_ABS:
    case F_MIN_RL: case F_MAX_RL: case F_POW_RL: case F_ATAN2_RL:
        return 1;
    case F_MIN_RR: case F_MAX_RR: case F_MUL_RR: case F_DIV_RR: case F_POW_RR: case F_ATAN2_RR:
        return 2;
    case F_DIV_LR:
        return 3;
    default:
        return 0;
    }
}

}  // namespace

// Algorithmic f32 operation counts (one op = one f32 +,-,*,/,sqrt,min,max result of the
// reference formula; compares, selects, sign flips and NaN tests count zero).  fix() adds 2.
uint32_t fwd_cost(uint32_t op, uint32_t side) {
    if (side == 3) return 2;
    uint32_t c;
    switch (op) {
    case Op::ADD: case Op::SUB: c = 2; break;
    case Op::MUL: c = side == 0 ? 10 : 2; break;
    case Op::DIV: c = 10; break;
    case Op::SQR: case Op::SQRT: c = 4; break;
    case Op::MIN: case Op::MAX: c = 2; break;
    case Op::ABS: c = 1; break;
    case Op::EXP: c = 2; break;
    case Op::LOG: c = 4; break;
    case Op::SIN: c = 16; break;
    case Op::COS: c = 17; break;
    case Op::POW: c = 8; break;
    case Op::MOD: c = 22; break;
    case Op::ATAN2: c = 8; break;
    default: c = 0; break;
    }
    return c + 2;
}

// Reverse: rule cost + 14 per receiving argument (fix 2 + MUL_II 10 + ADD_II 2).
uint32_t rev_cost(uint32_t op, uint32_t side) {
    if (side == 3) return 0;
    uint32_t nargs = side == 0 ? 2 : 1, rr;
    switch (op) {
    case Op::SQR: rr = 2; break;
    case Op::SQRT: rr = 17; break;
    case Op::DIV: rr = 24; break;
    case Op::EXP: rr = 2; break;
    case Op::LOG: rr = 10; break;
    case Op::SIN: rr = 17; break;
    case Op::COS: rr = 16; break;
    case Op::POW: rr = 24; break;
    case Op::ATAN2: rr = 28; break;
    default: rr = 0; break;
    }
    return rr + 14 * nargs;
}

int lower(const tsi_gpu_instruction *tape, uint32_t n, std::vector<FwdIns> &fwd, std::vector<RevIns> &rev,
          Lowered &L) {
    if (!tape || n < 5) return TSI_E_ARG;
    if (n > 4096) return TSI_E_TAPE_TOO_LONG;
    std::vector<int> last(n, -1);
    // Pruning class: a decided MIN/MAX stays decided in a child cell when every rule on the
    // way to its operands is inclusion-monotone in binary32 (child operands inside the
    // parent's).  Rounded +,-,*,/lit,sqrt,min,max,abs,floor are; the polynomial sin, cos,
    // exp and log routines are up to ulp-level wiggles; interval division by an interval,
    // pow, mod and atan2 are not.
    int prune_class = kPruneExact;
    for (uint32_t j = 4; j < n; ++j) {
        const auto &g = tape[j];
        uint32_t side = side_of(g), op = op_of(g);
        if ((side == 0 || side == 1) && larg_of(g) >= j) return TSI_E_ARG;
        if ((side == 0 || side == 2) && rarg_of(g) >= j) return TSI_E_ARG;
        if (side == 0 || side == 1) last[larg_of(g)] = std::max(last[larg_of(g)], int(j));
        if (side == 0 || side == 2) last[rarg_of(g)] = std::max(last[rarg_of(g)], int(j));
        if (side == 3) continue;   // lit/lit evaluates to [0, 0] whatever the opcode
        switch (op) {
        case Op::ADD: case Op::SUB: case Op::MUL: case Op::SQR: case Op::SQRT: case Op::NEG: case Op::MIN:
        case Op::MAX: case Op::ABS: case Op::RET: case Op::FLOOR:
            break;
        case Op::DIV:
            if (side != 1) prune_class = kPruneOff;
            break;
        case Op::SIN: case Op::COS: case Op::EXP: case Op::LOG:
            if (prune_class == kPruneExact) prune_class = kPruneApprox;
            break;
        default:
            prune_class = kPruneOff;
        }
    }
    fwd.clear();
    rev.clear();
    L = Lowered{};
    L.n = n;
    L.prune_class = prune_class;

    // ---- forward slot allocation ----
    std::vector<uint8_t> vslot(n, kNoSlot);
    std::vector<uint8_t> freelist;
    for (int s = kSlots - 1; s >= 4; --s) freelist.push_back(uint8_t(s));
    for (uint32_t k = 0; k < 4; ++k) vslot[k] = uint8_t(k);
    auto release = [&](uint32_t k) {
        if (vslot[k] != kNoSlot) { freelist.push_back(vslot[k]); vslot[k] = kNoSlot; }
    };
    for (uint32_t k = 0; k < 4; ++k)
        if (last[k] < 4) release(k);
    uint32_t live = 4, maxlive = 4;
    uint32_t pos = 0;
    for (uint32_t j = 4; j < n; ++j) {
        const auto &g = tape[j];
        uint32_t side = side_of(g), op = op_of(g), a = larg_of(g), b = rarg_of(g);
        FwdIns f{};
        f.form = form_of(op, side);
        f.a = (side == 0 || side == 1) ? vslot[a] : uint8_t(0);
        f.b = (side == 0 || side == 2) ? vslot[b] : uint8_t(0);
        if (f.a == kNoSlot || f.b == kNoSlot) return TSI_E_ARG;
        f.lit = g.literal;
        f.index = uint16_t(j - 4);
        f.modes = 0xFFFFu;
        f.cost = uint8_t(fwd_cost(op, side));
        uint8_t pk = push_kind(Form(f.form));
        f.npush_side = uint8_t(pk | (side << 2));
        if (op == Op::MIN || op == Op::MAX) {
            if (f.form == F_MIN_RR || f.form == F_MAX_RR || f.form == F_MIN_RL || f.form == F_MAX_RL) ++L.nminmax;
        }
        f.push = uint16_t(pos);
        pos += pk == 2 ? 2 : (pk ? 1 : 0);
        // operands whose last use is j die here
        if ((side == 0 || side == 1) && last[a] == int(j)) { release(a); --live; }
        if (side == 0 && b != a && last[b] == int(j)) { release(b); --live; }
        if (side == 2 && last[b] == int(j)) { release(b); --live; }
        if (freelist.empty()) return TSI_E_UNSUPPORTED;
        uint8_t d = freelist.back();
        freelist.pop_back();
        vslot[j] = d;
        f.d = d;
        ++live;
        maxlive = std::max(maxlive, live);
        if (last[j] < 0 && j != n - 1) { release(j); --live; }
        fwd.push_back(f);
        L.fwd_cost += fwd_cost(op, side);
    }
    if (pos > 0xFFFFu) return TSI_E_UNSUPPORTED;
    L.out_slot = vslot[n - 1];
    L.stack = pos;
    L.nfwd = uint32_t(fwd.size());

    // ---- reverse slot allocation (j descending) ----
    std::vector<uint8_t> sslot(n, kNoSlot);
    freelist.clear();
    for (int s = kSlots - 1; s >= 4; --s) freelist.push_back(uint8_t(s));
    for (uint32_t k = 0; k < 4; ++k) sslot[k] = uint8_t(k);
    uint32_t slive = 4, smax = 4;
    for (uint32_t j = n - 1; j >= 4; --j) {
        const auto &g = tape[j];
        uint32_t side = side_of(g), op = op_of(g), a = larg_of(g), b = rarg_of(g);
        const FwdIns &f = fwd[j - 4];
        RevIns r{};
        r.form = f.form;
        r.side = uint8_t(side);
        r.pop = f.push;
        r.lit = g.literal;
        r.npush = f.npush_side & 3u;
        r.cost = uint8_t(rev_cost(op, side));
        r.modes = 0xFFFFu;
        r.sa = r.sb = 0;
        if (j == n - 1) {
            r.flags |= R_SJ_ONE;
            if (sslot[j] != kNoSlot) { freelist.push_back(sslot[j]); sslot[j] = kNoSlot; --slive; }
        } else if (sslot[j] != kNoSlot) {
            r.sj = sslot[j];
            freelist.push_back(sslot[j]);
            sslot[j] = kNoSlot;
            --slive;
        } else {
            r.flags |= R_SJ_ZERO;
        }
        auto touch = [&](uint32_t k, uint8_t &slot, uint8_t ft) -> bool {
            if (sslot[k] == kNoSlot) {
                if (freelist.empty()) return false;
                sslot[k] = freelist.back();
                freelist.pop_back();
                ++slive;
                smax = std::max(smax, slive);
                r.flags |= ft;
            }
            slot = sslot[k];
            return true;
        };
        // lit/lit records do nothing in reverse; they are kept so that reverse record r
        // belongs to forward record nfwd-1-r (pruned programs select both by index)
        if (side != 3) {
            bool ok = true;
            if (side == 0 || side == 1) ok = ok && touch(a, r.sa, R_FT_A);
            if (side == 0 || side == 2) ok = ok && touch(b, r.sb, R_FT_B);
            if (!ok) return TSI_E_UNSUPPORTED;
            if (side == 0 && a == b && (r.flags & R_FT_A)) r.flags |= R_FT_B;
            if (uses_of(tape[a]) == 1u) r.flags |= R_U1_A;
            if (uses_of(tape[b]) == 1u) r.flags |= R_U1_B;
        }
        rev.push_back(r);
        if (j == 4) break;
    }
    L.nrev = uint32_t(rev.size());
    L.slots = std::max(maxlive, smax);
    return TSI_OK;
}

}  // namespace tsi
