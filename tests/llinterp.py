"""Independent scalar interpreter of a scene's `@_Z1fffff` body (test infrastructure).

It reads the `.ll` text with its own regular expressions -- it shares no code with the
two tape builders under test (oracle/ir.c and the library's tsi_ir.cpp) -- and evaluates
f(x, y, z, t) in float64 at a point.  The scalar semantics are those the reference's
interval rules give a degenerate interval [v, v] (compute.wgsl line numbers):

  fadd/fsub/fmul/fdiv/fneg    IEEE arithmetic                       (2535-2609)
  sqrtf                       sqrt(max(x, 0))                       SQRT_I 2572
  logf                        log(max(x, 0))                        LOG_I 2584
  expf, sinf, cosf            exp, sin, cos                         2576, 2935, 2959
  llvm.minnum / llvm.maxnum   min / max                             2613-2625
  llvm.fabs, llvm.floor       abs, floor                            2629, 2704
  fmodf(a, b)                 a - b * floor(a / b)  (not C's fmod)  MOD_II 2818
  powf(a, b)                  a ** rint(b), integer power           POW_II 2828
  atan2f(a, b)                atan2(b, a): ATAN2_II(larg, rarg)     2724, 3740
                              with y = rarg and x = larg; 0 when b == 0
  every result                clamped to [-1e6, 1e6]                fix 3462

`uses` counts register operand uses per SSA value the way the tape records them
(ir.rs:311-323, a value used twice by one instruction counts twice unless the
instruction is `fmul %a, %a`, which becomes SQR with one use, ir.rs:305-309).
"""
from __future__ import annotations

import math
import re
import struct

import numpy as np

BIG = 1.0e6

_DEF = re.compile(r"define[^\n]*@_Z1fffff\([^\n]*\{\n")
_ASSIGN = re.compile(r"^\s*%(\d+)\s*=\s*(.*)$")
_BIN = re.compile(r"^(fadd|fsub|fmul|fdiv)\s+(?:\w+\s+)*float\s+([^,\s]+)\s*,\s*([^\s,]+)")
_NEG = re.compile(r"^fneg\s+(?:\w+\s+)*float\s+([^\s,]+)")
_CALL = re.compile(r"^(?:tail\s+)?call\s+(?:\w+\s+)*float\s+@([\w.]+)\((.*?)\)")
_RET = re.compile(r"^\s*ret\s+float\s+([^\s,]+)")

# opcode numbering of ir.rs:8-43, for the histogram the fixtures hold
OPS = {"ADD": 5, "SUB": 6, "MUL": 7, "DIV": 8, "EXP": 9, "LOG": 10, "SQR": 11, "SQRT": 12, "POW": 13,
       "SIN": 14, "COS": 15, "NEG": 16, "MIN": 17, "MAX": 18, "ABS": 19, "FLOOR": 25, "RET": 28, "MOD": 29,
       "ATAN2": 30}

_CALLS = {"sqrtf": "SQRT", "llvm.sqrt.f32": "SQRT", "logf": "LOG", "expf": "EXP", "sinf": "SIN", "cosf": "COS",
          "llvm.minnum.f32": "MIN", "llvm.maxnum.f32": "MAX", "llvm.fabs.f32": "ABS", "llvm.floor.f32": "FLOOR",
          "powf": "POW", "fmodf": "MOD", "atan2f": "ATAN2"}


def _literal(tok: str) -> float:
    tok = tok.strip()
    if tok.startswith("0x"):
        (d,) = struct.unpack("<d", struct.pack("<Q", int(tok[2:], 16)))
        return float(np.float32(d))
    return float(np.float32(float(tok)))


class Program:
    """The body as a list of (dest, opname, operands); operands are ('r', ssa) or ('l', value)."""

    def __init__(self, text: str):
        m = _DEF.search(text)
        if not m:
            raise ValueError("no @_Z1fffff definition")
        body = text[m.end():text.index("\n}", m.end())]
        self.ins: list[tuple[int, str, list]] = []
        self.ret = None
        for line in body.split("\n"):
            if not line.strip():
                continue
            r = _RET.match(line)
            if r:
                self.ret = self._opnd(r.group(1))
                continue
            a = _ASSIGN.match(line)
            if not a:
                raise ValueError(f"unexpected line {line!r}")
            dest, rhs = int(a.group(1)), a.group(2)
            if (b := _BIN.match(rhs)):
                op = {"fadd": "ADD", "fsub": "SUB", "fmul": "MUL", "fdiv": "DIV"}[b.group(1)]
                ops = [self._opnd(b.group(2)), self._opnd(b.group(3))]
            elif (n := _NEG.match(rhs)):
                op, ops = "NEG", [self._opnd(n.group(1))]
            elif (c := _CALL.match(rhs)):
                op = _CALLS[c.group(1)]
                ops = [self._opnd(x.strip().split()[-1]) for x in c.group(2).split(",")]
            else:
                raise ValueError(f"unsupported instruction {rhs!r}")
            self.ins.append((dest, op, ops))
        if self.ret is None:
            raise ValueError("no ret")

    @staticmethod
    def _opnd(tok: str):
        tok = tok.strip().rstrip(",")
        if tok.startswith("%"):
            return ("r", int(tok[1:]))
        return ("l", _literal(tok))

    # ---- structure -------------------------------------------------------------------
    def histogram(self) -> dict:
        """Opcode counts of the tape body as the reference's builder classifies it."""
        h: dict[str, int] = {}
        for _, op, ops in self.ins:
            if op == "MUL" and ops[0] == ops[1]:
                op = "SQR"
            h[op] = h.get(op, 0) + 1
        h["RET"] = 1
        return h

    def uses(self) -> dict:
        """ssa -> number of register uses, ret included."""
        u: dict[int, int] = {}
        for _, op, ops in self.ins:
            if op == "MUL" and ops[0] == ops[1]:
                ops = ops[:1]
            for kind, v in ops:
                if kind == "r":
                    u[v] = u.get(v, 0) + 1
        if self.ret[0] == "r":
            u[self.ret[1]] = u.get(self.ret[1], 0) + 1
        return u

    # ---- evaluation --------------------------------------------------------------------
    def eval(self, x: float, y: float, z: float, t: float):
        """Returns (value, margin): margin = distance of the closest floor/mod quotient to an
        integer (float32 and float64 may round differently there) or of atan2's y to 0."""
        env = {0: x, 1: y, 2: z, 3: t}
        margin = math.inf

        def val(o):
            return env[o[1]] if o[0] == "r" else o[1]

        for dest, op, ops in self.ins:
            a = val(ops[0])
            b = val(ops[1]) if len(ops) > 1 else None
            if op == "ADD":
                r = a + b
            elif op == "SUB":
                r = a - b
            elif op == "MUL":
                r = a * b
            elif op == "DIV":
                r = a / b if b != 0.0 else math.copysign(math.inf, a) * math.copysign(1.0, b)
            elif op == "NEG":
                r = -a
            elif op == "SQRT":
                r = math.sqrt(max(a, 0.0))
            elif op == "LOG":
                r = math.log(a) if a > 0.0 else -math.inf
            elif op == "EXP":
                r = math.exp(a) if a < 700.0 else math.inf
            elif op == "SIN":
                r = math.sin(a)
            elif op == "COS":
                r = math.cos(a)
            elif op == "MIN":
                r = min(a, b)
            elif op == "MAX":
                r = max(a, b)
            elif op == "ABS":
                r = abs(a)
            elif op == "FLOOR":
                margin = min(margin, abs(a - round(a)))
                r = math.floor(a)
            elif op == "MOD":
                q = a / b
                margin = min(margin, abs(q - round(q)))
                r = a - b * math.floor(q)
            elif op == "POW":
                r = a ** int(round(b))
            elif op == "ATAN2":
                margin = min(margin, abs(b))
                r = math.atan2(b, a) if b != 0.0 else 0.0
            else:
                raise AssertionError(op)
            if math.isnan(r):
                r = 0.0
                margin = 0.0
            env[dest] = min(max(r, -BIG), BIG)
        return val(self.ret), margin

    def eval_many(self, x, y, z, t) -> np.ndarray:
        """eval() over arrays of points at once (float64; same semantics, no margin), for the
        dense sampling of large octrees.  Supports the opcodes of the seeded scene generators."""
        x, y, z, t = (np.asarray(a, np.float64) for a in np.broadcast_arrays(x, y, z, t))
        env = {0: x, 1: y, 2: z, 3: t}

        def val(o):
            return env[o[1]] if o[0] == "r" else o[1]

        with np.errstate(all="ignore"):
            for dest, op, ops in self.ins:
                a = val(ops[0])
                b = val(ops[1]) if len(ops) > 1 else None
                if op == "ADD":
                    r = a + b
                elif op == "SUB":
                    r = a - b
                elif op == "MUL":
                    r = a * b
                elif op == "DIV":
                    r = np.divide(a, b)
                elif op == "NEG":
                    r = -a
                elif op == "SQRT":
                    r = np.sqrt(np.maximum(a, 0.0))
                elif op == "EXP":
                    r = np.exp(np.minimum(a, 700.0))
                elif op == "SIN":
                    r = np.sin(a)
                elif op == "COS":
                    r = np.cos(a)
                elif op == "MIN":
                    r = np.minimum(a, b)
                elif op == "MAX":
                    r = np.maximum(a, b)
                elif op == "ABS":
                    r = np.abs(a)
                elif op == "FLOOR":
                    r = np.floor(a)
                else:
                    raise NotImplementedError(op)
                r = np.asarray(r, np.float64)
                np.copyto(r, 0.0, where=np.isnan(r))
                env[dest] = np.clip(r, -BIG, BIG, out=r if r.shape == x.shape else None)
        return np.broadcast_to(val(self.ret), x.shape).astype(np.float64)
