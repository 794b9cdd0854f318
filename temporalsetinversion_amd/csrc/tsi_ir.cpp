// tsi_ir.cpp -- host tape builder: .ll text -> 32-byte GPUInstruction records.
//
// Restates src/ir.rs (IR::new 334-366, parseLLVMIR 190-331, Instruction::to_gpu 139-181)
// so that a host without the Rust builder produces byte-identical tapes:
//   * the body is the text between the first newline after the first "fffff" and the first
//     '}' after it (ir.rs:191-196); its first (empty) line is skipped;
//   * opcode by token: ret, fadd/fneg/fsub/fmul/fdiv, or for `tail call` the first of
//     sqrt, log, exp, min, max, floor, abs, sin, cos, pow, mod, atan2 contained in the
//     callee token (ir.rs:204-235);
//   * operands: %k registers (k > 3 renumbered k-1, ir.rs:297-307), LLVM double-hex
//     literals rounded to f32, decimal literals parsed as f32 (ir.rs:238-293);
//   * MUL x,x becomes SQR (ir.rs:310-313); uses are counted per register operand and the
//     last instruction gets one extra use (ir.rs:315-327, 355);
//   * packing: opcode_uses = (op | side<<14)<<16 | uses, larg_rarg = larg<<16 | rarg, the
//     literal is the right operand's when both are literals (ir.rs:140-181).
// Inputs ir.rs would panic on are reported as TSI_E_PARSE.
#include "tsi.h"
#include "tsi_tape.h"

#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <string>
#include <string_view>
#include <vector>

namespace {

namespace Op = tsi::Op;

struct Operand {
    bool reg = false;
    uint32_t index = 0;
    float value = 0.0f;
    bool operator==(const Operand &o) const {
        return reg == o.reg && (reg ? index == o.index : value == o.value);
    }
};

struct Builder {
    uint32_t op = 0;
    uint32_t uses = 0;
    Operand l, r;
    float val[2] = {0.0f, 0.0f};
};

bool unary(uint32_t op) {
    switch (op) {
    case Op::NOP: case Op::ARGX: case Op::ARGY: case Op::ARGZ: case Op::LOG: case Op::SQR: case Op::SQRT:
    case Op::SIN: case Op::COS: case Op::NEG: case Op::EXP: case Op::TRUNC: case Op::FLOOR: case Op::CEIL:
    case Op::SIGN: case Op::ABS: case Op::RET:
        return true;
    default:
        return false;
    }
}

bool tail(uint32_t op) {
    switch (op) {
    case Op::EXP: case Op::LOG: case Op::SQRT: case Op::MIN: case Op::MAX: case Op::FLOOR: case Op::CEIL:
    case Op::ABS: case Op::SIN: case Op::COS: case Op::POW: case Op::MOD: case Op::ATAN2:
        return true;
    default:
        return false;
    }
}

bool ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

std::vector<std::string_view> tokens(std::string_view line) {
    std::vector<std::string_view> t;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && ws(line[i])) ++i;
        size_t s = i;
        while (i < line.size() && !ws(line[i])) ++i;
        if (i > s) t.push_back(line.substr(s, i - s));
    }
    return t;
}

// token[a .. len-b]
bool cut(std::string_view t, size_t a, size_t b, std::string_view &out) {
    if (a + b > t.size()) return false;
    out = t.substr(a, t.size() - a - b);
    return !out.empty();
}

bool as_index(std::string_view s, uint32_t &v) {
    if (!s.empty() && s[0] == '+') s.remove_prefix(1);
    if (s.empty() || s.size() > 10) return false;
    uint64_t x = 0;
    for (char c : s) {
        if (c < '0' || c > '9') return false;
        x = x * 10 + uint64_t(c - '0');
    }
    if (x > 0xFFFFFFFFull) return false;
    v = uint32_t(x);
    return true;
}

bool as_decimal(std::string_view s, float &v) {
    if (s.empty() || s.size() > 100) return false;
    for (char c : s)
        if (!((c >= '0' && c <= '9') || c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-')) return false;
    std::string z(s);
    char *end = nullptr;
    v = std::strtof(z.c_str(), &end);
    return end == z.c_str() + z.size();
}

bool as_hex_double(std::string_view s, float &v) {
    if (s.empty() || s.size() > 16) return false;
    uint64_t bits = 0;
    for (char c : s) {
        int d;
        if (c >= '0' && c <= '9') d = c - '0';
        else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
        else if (c >= 'A' && c <= 'F') d = c - 'A' + 10;
        else return false;
        bits = (bits << 4) | uint64_t(d);
    }
    double d;
    std::memcpy(&d, &bits, sizeof d);
    v = float(d);
    return true;
}

bool has(std::string_view t, const char *needle) { return t.find(needle) != std::string_view::npos; }

// Operand forms of ir.rs: register "%k" with `rtrim` trailing characters dropped; hex
// literal "0x..." with `htrim` dropped (allowed only where ir.rs accepts hex); decimal
// literal with `dtrim` dropped.
bool operand(std::string_view t, size_t rtrim, bool hex_ok, size_t htrim, size_t dtrim, Operand &o) {
    std::string_view s;
    if (has(t, "%")) {
        o.reg = true;
        return cut(t, 1, rtrim, s) && as_index(s, o.index);
    }
    o.reg = false;
    if (hex_ok && has(t, "0x")) return cut(t, 2, htrim, s) && as_hex_double(s, o.value);
    return cut(t, 0, dtrim, s) && as_decimal(s, o.value);
}

uint32_t classify_call(std::string_view callee) {
    static const struct { const char *k; uint32_t op; } order[] = {
        {"sqrt", Op::SQRT}, {"log", Op::LOG}, {"exp", Op::EXP}, {"min", Op::MIN}, {"max", Op::MAX},
        {"floor", Op::FLOOR}, {"abs", Op::ABS}, {"sin", Op::SIN}, {"cos", Op::COS}, {"pow", Op::POW},
        {"mod", Op::MOD}, {"atan2", Op::ATAN2}};
    for (const auto &e : order)
        if (has(callee, e.k)) return e.op;
    return Op::NOP;
}

int parse_line(std::string_view line, std::vector<Builder> &ins) {
    auto tk = tokens(line);
    if (tk.empty()) return TSI_E_PARSE;
    Builder b;
    const bool ret = tk[0] == "ret";
    if (ret) {
        b.op = Op::RET;
    } else {
        if (tk.size() < 3) return TSI_E_PARSE;
        std::string_view k = tk[2];
        if (k == "fadd") b.op = Op::ADD;
        else if (k == "fneg") b.op = Op::NEG;
        else if (k == "fsub") b.op = Op::SUB;
        else if (k == "fmul") b.op = Op::MUL;
        else if (k == "fdiv") b.op = Op::DIV;
        else if (k == "tail") {
            if (tk.size() < 6) return TSI_E_PARSE;
            b.op = classify_call(tk[5]);
        } else {
            b.op = Op::NOP;
        }
    }
    bool ok;
    if (tail(b.op) && unary(b.op)) {
        if (tk.size() < 7) return TSI_E_PARSE;
        ok = operand(tk[6], 1, true, 1, 0, b.l);
    } else if (tail(b.op)) {
        if (tk.size() < 9) return TSI_E_PARSE;
        ok = operand(tk[6], 1, false, 0, 0, b.l) && operand(tk[8], 1, true, 1, 1, b.r);
    } else if (unary(b.op)) {
        if (ret) {
            std::string_view s;
            if (tk.size() < 3) return TSI_E_PARSE;
            b.l.reg = true;
            ok = cut(tk[2], 1, 0, s) && as_index(s, b.l.index);
        } else {
            if (tk.size() < 5) return TSI_E_PARSE;
            ok = operand(tk[4], 0, false, 0, 0, b.l);
        }
    } else {
        if (tk.size() < 6) return TSI_E_PARSE;
        ok = operand(tk[4], 1, true, 1, 1, b.l) && operand(tk[5], 0, true, 0, 0, b.r);
    }
    if (!ok) return TSI_E_PARSE;
    for (Operand *o : {&b.l, &b.r})
        if (o->reg && o->index > 3) o->index -= 1;
    if (b.op == Op::MUL && b.l == b.r) {
        b.op = Op::SQR;
        b.r = Operand{};
    }
    for (Operand *o : {&b.l, &b.r}) {
        if (!o->reg) continue;
        if (o->index >= ins.size()) return TSI_E_PARSE;
        ins[o->index].uses++;
    }
    ins.push_back(b);
    return TSI_OK;
}

tsi_gpu_instruction pack(const Builder &b, uint32_t index) {
    tsi_gpu_instruction g{};
    uint32_t side = 3, larg = 0, rarg = 0;
    float literal = 0.0f;
    if (b.l.reg) { side -= 2; larg = b.l.index & 0xFFFFu; } else { literal = b.l.value; }
    if (b.r.reg) { side -= 1; rarg = b.r.index & 0xFFFFu; } else { literal = b.r.value; }
    g.index_ = index & 0xFFFFu;
    g.opcode_uses = (((b.op | (side << 14)) & 0xFFFFu) << 16) | (b.uses & 0xFFFFu);
    g.larg_rarg = (larg << 16) | rarg;
    g.literal = literal;
    g.val[0] = b.val[0];
    g.val[1] = b.val[1];
    return g;
}

}  // namespace

extern "C" int tsi_tape_from_ll(const char *text, size_t len, float time, tsi_gpu_instruction *out, uint32_t cap,
                                uint32_t *n) {
    if (!text) return TSI_E_ARG;
    try {
        std::string_view all(text, len);
        size_t a = all.find("fffff");
        if (a == std::string_view::npos) return TSI_E_PARSE;
        std::string_view b = all.substr(a);
        size_t start = b.find('\n'), end = b.find('}');
        if (start == std::string_view::npos || end == std::string_view::npos || end <= start) return TSI_E_PARSE;
        std::string_view body = b.substr(start, end - 1 - start);

        std::vector<Builder> ins(4);
        for (uint32_t k = 0; k < 4; ++k) ins[k].op = Op::ARGX + k;
        for (int k = 0; k < 3; ++k) { ins[k].val[0] = -5.0f; ins[k].val[1] = 5.0f; }
        ins[3].val[0] = ins[3].val[1] = time;

        size_t pos = 0;
        bool first = true;
        while (true) {
            size_t nl = body.find('\n', pos);
            std::string_view line = body.substr(pos, nl == std::string_view::npos ? std::string_view::npos : nl - pos);
            if (!first) {
                int rc = parse_line(line, ins);
                if (rc != TSI_OK) return rc;
            }
            first = false;
            if (nl == std::string_view::npos) break;
            pos = nl + 1;
        }
        ins.back().uses++;
        if (n) *n = uint32_t(ins.size());
        if (out) {
            uint32_t m = uint32_t(ins.size() < 4096 ? ins.size() : 4096);
            if (m > cap) m = cap;
            for (uint32_t i = 0; i < m; ++i) out[i] = pack(ins[i], i);
        }
        return ins.size() > 4096 ? TSI_E_TAPE_TOO_LONG : TSI_OK;
    } catch (...) {
        return TSI_E_OOM;
    }
}
