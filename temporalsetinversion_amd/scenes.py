"""Seeded scene generators that emit straight-line `.ll` text.

The reference's scenes are GLSL-like C++ fields f(x, y, z, t) compiled with
`clang -S -emit-llvm -O3` (scenes/compile.sh:2); the tape builder then reads the
body of `@_Z1fffff` (src/ir.rs:190-331).  The generators below write the same
dialect directly (one SSA value per line, `%0..%3` = x, y, z, t, values from `%5`
on, constants as LLVM double-hex literals, commutative constants on the right)
so that generated scenes go through exactly the same parser as hand-written ones.

Workloads (seeded, numpy PCG64):
  spheres   -- min-union of 8 spheres, centres c0 + 0.2 sin(2 pi t + phi)
  spheres_linear -- the same spheres moving linearly, c0 + v t (no sin: bit-exact op set)
  capsules  -- smooth-min (k = 0.05) union of 64 capsules translating linearly in t,
               under a twist warp of angle 2 t z
  csg       -- 256 spheres / boxes / tori combined by min / max / smooth-min under
               sin/cos domain warps of amplitude 0.05
"""
from __future__ import annotations

import math
import struct

import numpy as np


def _hexlit(v: float) -> str:
    """LLVM's spelling of a float constant: the double with the f32's value, in hex."""
    f = float(np.float32(v))
    return "0x%016X" % struct.unpack("<Q", struct.pack("<d", f))[0]


class Tape:
    """Minimal SSA emitter for the `@_Z1fffff` body."""

    def __init__(self):
        self.lines: list[str] = []
        self.next = 5          # %4 is the entry label in clang output
        self.x, self.y, self.z, self.t = "%0", "%1", "%2", "%3"

    def _new(self, rhs: str) -> str:
        r = f"%{self.next}"
        self.next += 1
        self.lines.append(f"  {r} = {rhs}")
        return r

    @staticmethod
    def _opnd(a) -> str:
        return a if isinstance(a, str) else _hexlit(a)

    def _bin(self, op, a, b):
        if not isinstance(a, str) and not isinstance(b, str):
            raise ValueError("constant folding is the caller's job")
        if op in ("fadd", "fmul") and not isinstance(a, str):
            a, b = b, a
        return self._new(f"{op} float {self._opnd(a)}, {self._opnd(b)}")

    def add(self, a, b): return self._bin("fadd", a, b)
    def sub(self, a, b): return self._bin("fsub", a, b)
    def mul(self, a, b): return self._bin("fmul", a, b)
    def div(self, a, b): return self._bin("fdiv", a, b)
    def neg(self, a): return self._new(f"fneg float {a}")

    def _call1(self, fn, a):
        return self._new(f"tail call float @{fn}(float {self._opnd(a)}) #10")

    def _call2(self, fn, a, b):
        if not isinstance(a, str):
            a, b = b, a
        return self._new(f"tail call float @{fn}(float {a}, float {self._opnd(b)}) #10")

    def sqrt(self, a): return self._call1("sqrtf", a)
    def abs(self, a): return self._call1("llvm.fabs.f32", a)
    def sin(self, a): return self._call1("sinf", a)
    def cos(self, a): return self._call1("cosf", a)
    def exp(self, a): return self._call1("expf", a)
    def min(self, a, b): return self._call2("llvm.minnum.f32", a, b)
    def max(self, a, b): return self._call2("llvm.maxnum.f32", a, b)

    def sq(self, a): return self.mul(a, a)

    def length3(self, a, b, c):
        return self.sqrt(self.add(self.add(self.sq(a), self.sq(b)), self.sq(c)))

    def length2(self, a, b):
        return self.sqrt(self.add(self.sq(a), self.sq(b)))

    def smin(self, a, b, k):
        """min(a, b) - max(k - |a - b|, 0)^2 * 0.25 / k (scenes/sponza.cpp:66-69)."""
        h = self.max(self.sub(k, self.abs(self.sub(a, b))), 0.0)
        return self.sub(self.min(a, b), self.mul(self.sq(h), 0.25 / k))

    def to_ll(self, result: str, name: str = "scene") -> str:
        body = "\n".join(self.lines + [f"  ret float {result}"])
        return (f"; ModuleID = '{name}.cpp'\nsource_filename = \"{name}.cpp\"\n\n"
                "; Function Attrs: nounwind uwtable\n"
                "define dso_local float @_Z1fffff(float %0, float %1, float %2, float %3) "
                "local_unnamed_addr #5 {\n" + body + "\n}\n\n"
                "declare float @sqrtf(float) local_unnamed_addr #8\n")


def spheres(seed: int = 0, n: int = 8, linear: bool = False) -> str:
    rng = np.random.default_rng(seed)
    T = Tape()
    r = rng.uniform(0.1, 0.25, n)
    c0 = rng.uniform(-0.6, 0.6, (n, 3))
    phi = rng.uniform(0.0, 2 * math.pi, n)
    v = rng.uniform(-0.3, 0.3, (n, 3))
    w = T.mul(T.t, 2 * math.pi) if not linear else None
    d = None
    for k in range(n):
        if linear:
            cx = [T.add(T.mul(T.t, float(v[k, a])), float(c0[k, a])) for a in range(3)]
        else:
            s = T.mul(T.sin(T.add(w, float(phi[k]))), 0.2)
            cx = [T.add(s, float(c0[k, a])) for a in range(3)]
        dist = T.sub(T.length3(T.sub(T.x, cx[0]), T.sub(T.y, cx[1]), T.sub(T.z, cx[2])), float(r[k]))
        d = dist if d is None else T.min(d, dist)
    return T.to_ll(d, "spheres_linear" if linear else "spheres")


def capsules(seed: int = 1, n: int = 64, k: float = 0.05) -> str:
    rng = np.random.default_rng(seed)
    T = Tape()
    # twist about z by angle 2 t z
    th = T.mul(T.mul(T.t, 2.0), T.z)
    c, s = T.cos(th), T.sin(th)
    px = T.sub(T.mul(T.x, c), T.mul(T.y, s))
    py = T.add(T.mul(T.x, s), T.mul(T.y, c))
    pz = T.z
    d = None
    for i in range(n):
        a0 = rng.uniform(-0.9, 0.9, 3)
        length = rng.uniform(0.1, 0.5)
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        ba = axis * length
        rad = rng.uniform(0.02, 0.08)
        vel = rng.uniform(-0.4, 0.4, 3)
        inv = 1.0 / float(np.dot(ba, ba))
        # a(t) = a0 + vel t ; pa = p - a(t)
        pa = [T.sub(T.sub(p, float(a0[j])), T.mul(T.t, float(vel[j]))) for j, p in enumerate((px, py, pz))]
        dot = T.add(T.add(T.mul(pa[0], float(ba[0] * inv)), T.mul(pa[1], float(ba[1] * inv))),
                    T.mul(pa[2], float(ba[2] * inv)))
        h = T.min(T.max(dot, 0.0), 1.0)
        q = [T.sub(pa[j], T.mul(h, float(ba[j]))) for j in range(3)]
        dist = T.sub(T.length3(*q), float(rad))
        d = dist if d is None else T.smin(d, dist, k)
    return T.to_ll(d, "capsules")


def csg(seed: int = 2, n: int = 256, amp: float = 0.05) -> str:
    rng = np.random.default_rng(seed)
    T = Tape()
    # domain warp: p += amp * sin(f . (y, z, x) + w t)  (sin/cos in x, y, z, t)
    f = rng.uniform(1.0, 6.0, (3, 2))
    x = T.add(T.x, T.mul(T.sin(T.add(T.mul(T.y, float(f[0, 0])), T.mul(T.t, float(f[0, 1])))), amp))
    y = T.add(T.y, T.mul(T.cos(T.add(T.mul(T.z, float(f[1, 0])), T.mul(T.t, float(f[1, 1])))), amp))
    z = T.add(T.z, T.mul(T.sin(T.add(T.mul(T.x, float(f[2, 0])), T.mul(T.t, float(f[2, 1])))), amp))
    d = None
    for i in range(n):
        c = rng.uniform(-1.2, 1.2, 3)
        size = rng.uniform(0.02, 0.2)
        kind = rng.integers(0, 3)
        q = [T.sub(x, float(c[0])), T.sub(y, float(c[1])), T.sub(z, float(c[2]))]
        if kind == 0:      # sphere
            prim = T.sub(T.length3(*q), float(size))
        elif kind == 1:    # box, Chebyshev form: max_i |q_i| - b_i
            b = rng.uniform(0.5, 1.0, 3) * size
            e = [T.sub(T.abs(q[j]), float(b[j])) for j in range(3)]
            prim = T.max(T.max(e[0], e[1]), e[2])
        else:              # torus in the xz plane
            R = size
            rr = size * float(rng.uniform(0.2, 0.5))
            prim = T.sub(T.length2(T.sub(T.length2(q[0], q[2]), R), q[1]), rr)
        if d is None:
            d = prim
        else:
            op = rng.integers(0, 8)
            if op < 5:
                d = T.min(d, prim)
            elif op < 7:
                d = T.smin(d, prim, float(rng.uniform(0.02, 0.08)))
            else:
                d = T.max(d, T.neg(prim)) if rng.integers(0, 2) else T.min(d, prim)
    return T.to_ll(d, "csg")


def sphere_unit(radius: float = 1.0) -> str:
    """|p| - r, a static single sphere (smallest useful tape)."""
    T = Tape()
    return T.to_ll(T.sub(T.length3(T.x, T.y, T.z), radius), "sphere")


def mixed(seed: int = 0, n: int = 12) -> str:
    """A seeded union/intersection/difference of n primitives built from most of the forms the
    evaluators distinguish (the generator of the randomised parity tests): spheres and boxes,
    planes with a literal on the left (c - x), products of two registers, divisions by
    1 + q^2, MIN/MAX against literals, a square root of |u v|, and a sine of a value the primitive
    uses twice (a transcendental record whose operand has two uses: the reverse pass cannot
    short-circuit it).  Centres move with t."""
    rng = np.random.default_rng(seed)
    T = Tape()
    d = None
    for i in range(n):
        c = rng.uniform(-1.0, 1.0, 3)
        w = float(rng.uniform(0.5, 3.0))
        ph = float(rng.uniform(0.0, 6.283))
        amp = float(rng.uniform(0.0, 0.25))
        # moving centre: c_x(t) = c0 + amp sin(w t + ph)
        cx = T.add(T.mul(T.sin(T.add(T.mul(T.t, w), ph)), amp), float(c[0]))
        q = [T.sub(T.x, cx), T.sub(T.y, float(c[1])), T.sub(T.z, float(c[2]))]
        size = float(rng.uniform(0.1, 0.4))
        kind = int(rng.integers(0, 7))
        if kind == 0:        # sphere
            prim = T.sub(T.length3(*q), size)
        elif kind == 1:      # box
            e = [T.sub(T.abs(q[j]), size * float(rng.uniform(0.5, 1.0))) for j in range(3)]
            prim = T.max(T.max(e[0], e[1]), e[2])
        elif kind == 2:      # slab between two planes: c - (a x + b y), then clamped
            pl = T.add(T.mul(q[0], float(rng.uniform(0.2, 1.0))), T.mul(q[1], float(rng.uniform(-1.0, 1.0))))
            prim = T.max(T.sub(size, pl), T.sub(pl, 2.0 * size))
        elif kind == 3:      # saddle: q0 q1 - size, a product of two registers
            prim = T.sub(T.mul(q[0], q[1]), T.mul(q[2], size))
        elif kind == 4:      # bump: c / (1 + q0^2) against the height, min/max with literals
            bump = T.div(size, T.add(T.sq(q[0]), 1.0))
            prim = T.max(T.min(T.sub(q[1], bump), 0.5), -0.5)
        elif kind == 5:      # sqrt(|q0 q2|) + |q1| - size
            prim = T.sub(T.add(T.sqrt(T.abs(T.mul(q[0], q[2]))), T.abs(q[1])), size)
        else:                # rippled sphere: |q| - size + 0.05 sin(s) with s used twice
            s_ = T.mul(q[1], float(rng.uniform(2.0, 8.0)))
            r = T.add(T.mul(T.sin(s_), 0.05), T.mul(s_, 0.01))
            prim = T.add(T.sub(T.length3(*q), size), r)
        if d is None:
            d = prim
        else:
            op = int(rng.integers(0, 6))
            if op < 3:
                d = T.min(d, prim)
            elif op == 3:
                d = T.smin(d, prim, float(rng.uniform(0.02, 0.1)))
            elif op == 4:
                d = T.max(d, T.neg(prim))
            else:
                d = T.max(d, T.sub(prim, 0.6))
    return T.to_ll(d, "mixed")


def mixed_scenes() -> list:
    """(seed, n) of the randomised parity tests' scenes (their kernels are compiled ahead by the build)."""
    return [(11, 6), (12, 10), (13, 14), (14, 20)]


SCENES = {
    "mixed": mixed,
    "spheres": spheres,
    "spheres_linear": lambda seed=0: spheres(seed, linear=True),
    "capsules": capsules,
    "csg": csg,
}


# ---- the batched and served workloads ----

FRAME_SIZES = [(512, 512), (640, 480), (800, 600), (1024, 768), (1024, 1024), (1280, 720), (1600, 900), (1920, 1080)]


def batch_tapes(count: int = 16, seed: int = 4) -> list:
    """Tapes of the batched workload: `count` seeded tapes alternating the moving-spheres and the
    capsules generators, sized for about 100-2000 tape records (a sphere takes ~17, a capsule ~37)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 2 == 0:
            out.append(spheres(seed=1000 + i, n=int(rng.integers(6, 118))))
        else:
            out.append(capsules(seed=1000 + i, n=int(rng.integers(3, 55))))
    return out


def served_tapes(count: int = 32, seed: int = 5) -> list:
    """Tapes of the served workload: `count` seeded capsule scenes of about 200-2000 records."""
    rng = np.random.default_rng(seed)
    return [capsules(seed=2000 + i, n=int(rng.integers(6, 54))) for i in range(count)]


def batch_requests(count: int = 256, ntapes: int = 16, seed: int = 6, levels=(6, 10), sizes=None) -> list:
    """Independent requests of the batched workload: tape ~ U{0..ntapes-1}, localsubdiv ~ U{levels},
    a swept t-window [t1 - w, t1] with w ~ U[1/30, 1/2] and t1 ~ U[w, 1], a frame size from
    512x512 to 1920x1080."""
    rng = np.random.default_rng(seed)
    sizes = sizes or FRAME_SIZES
    out = []
    for _ in range(count):
        w = float(np.float32(rng.uniform(1.0 / 30.0, 0.5)))
        t1 = float(np.float32(rng.uniform(w, 1.0)))
        W, H = sizes[int(rng.integers(0, len(sizes)))]
        out.append(dict(tape=int(rng.integers(0, ntapes)), localsubdiv=int(rng.integers(levels[0], levels[1] + 1)),
                        window=w, time=t1, width=W, height=H))
    return out


def batch_test_sets() -> dict:
    """Tape sets of the batch parity tests (their kernels are compiled ahead by the build)."""
    return {"mixed4": [spheres(0), capsules(1, n=6), spheres(3, n=12, linear=True), csg(2, n=16)],
            "pair": [capsules(1, n=6), spheres(0)], "one": [spheres(0)]}
